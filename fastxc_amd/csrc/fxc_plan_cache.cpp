#include "fxc_plan_cache.hpp"

#include <map>
#include <memory>
#include <mutex>
#include <tuple>

namespace fxc {

namespace {
std::mutex g_mu;
std::map<std::tuple<int, int, int, int>, std::unique_ptr<FftPlan>> g_cplx;
std::map<std::tuple<int, int, int, int, bool>, std::unique_ptr<RealFftPlan>> g_real;
}  // namespace

int plan_cache_complex(int n, int sign, const FftPlan **out, int max_radix) {
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, n, sign, max_radix);
  auto it = g_cplx.find(key);
  if (it == g_cplx.end()) {
    std::unique_ptr<FftPlan> p(new FftPlan());
    FXC_TRY(fft_plan_create(p.get(), n, sign, max_radix));
    it = g_cplx.emplace(key, std::move(p)).first;
  }
  *out = it->second.get();
  return FXC_OK;
}

int plan_cache_real(int n, int sign, const RealFftPlan **out, int max_radix, bool largest_last) {
  int dev = 0;
  FXC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(dev, n, sign, max_radix, largest_last);
  auto it = g_real.find(key);
  if (it == g_real.end()) {
    std::unique_ptr<RealFftPlan> p(new RealFftPlan());
    FXC_TRY(real_fft_plan_create(p.get(), n, sign, max_radix, largest_last));
    it = g_real.emplace(key, std::move(p)).first;
  }
  *out = it->second.get();
  return FXC_OK;
}

void plan_cache_release() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_cplx) fft_plan_destroy(kv.second.get());
  for (auto &kv : g_real) real_fft_plan_destroy(kv.second.get());
  g_cplx.clear();
  g_real.clear();
}

}  // namespace fxc

extern "C" void fxc_release_plans(void) { fxc::plan_cache_release(); }
