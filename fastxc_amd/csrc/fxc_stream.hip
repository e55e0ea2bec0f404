// Streaming segment updates with device-resident per-pair stacks.
//
// The reference stacks offline: xc_fast writes every timestamp's
// correlations, the linear stack sums them in float64 in Python
// (fastxc/operators/stacking/linear_stack.py:228-247) and ncf_pws /
// ncf_tfpws re-read them (native/pws/io/sourcepack_io.cu:385-478).  This
// engine keeps that state in HBM instead and updates it as segments arrive:
//
//   submit(channels, traces)  ->  sac2spec of the segment's channels
//                              ->  xc of every channel pair (lower id = src)
//                              ->  acc[pair] += (double) ccf,  count[pair] += 1
//                                  (and the ccf kept in a per-pair history ring)
//
// Segments run in "lanes": each lane owns a HIP stream, a sac2spec plan, an
// xc plan and their workspaces (sized for the largest segment and warmed up
// at creation, so the steady state allocates nothing and never synchronises
// the device).  Segments that arrive while others are on the GPU are
// admitted as lanes free up, and their transfers and kernels overlap.  The
// state updates of all lanes are enqueued on one update stream in admission
// order, so a pair is never updated by two segments at once, each history
// position holds the newest record assigned to it, and the float64 sums add
// in the order a sequential run adds them.  A segment whose new pairs do not fit
// the state pool is refused (its ticket reports FXC_ERR_ALLOC).
//
// Read-out: the linear stack is acc / count rounded once to float (the
// reference's float64 running sum); PWS / TF-PWS are computed from the
// history ring in submission order with the reference's float prestack
// (sub-stacks of B records) and the fxc_pws / fxc_tfpws kernels.
#include "fxc_common.hpp"
#include "fastxc_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

int xc_set_pairs_on(fxc_xc_plan *p, int nsta, const int *h_src, const int *h_rec, long long npairs,
                    hipStream_t stream);
namespace fxc {
int stream_pws_close(float2 *S, float *open, const int *d_slot, const int *d_cnt, long long nclose,
                     int n, hipStream_t st, const float *rows = nullptr);
int stream_pws_read(const float2 *S, const float *open, const double *acc, const int *d_slot,
                    const long long *d_total, const int *d_nopen, long long npairs, int n,
                    float power, float *out, hipStream_t st);
}  // namespace fxc

namespace fxc {

constexpr int kStrThreads = 256;

// acc[slot[p]][l] += (double) cc[p][l]; the ccf also goes to history row
// hpos[p] of the slot.  A pair occurs once per segment, and segments are
// applied one after another, so no two threads update the same element.
__global__ void __launch_bounds__(kStrThreads)
stream_update_kernel(const float *__restrict__ cc, int ncc, long long npairs,
                     const int *__restrict__ slot, const int *__restrict__ hpos,
                     double *__restrict__ acc, float *__restrict__ hist, int history,
                     float *__restrict__ open) {
  const int l = blockIdx.x * kStrThreads + threadIdx.x;
  if (l >= ncc) return;
  for (long long p = blockIdx.y; p < npairs; p += gridDim.y) {
    const float v = cc[p * ncc + l];
    const long long s = slot[p];
    acc[s * ncc + l] += (double)v;
    if (history > 0) hist[(s * history + hpos[p]) * ncc + l] = v;
    if (open) open[s * ncc + l] += v;  // the open PWS sub-stack, float in record order
  }
}

// Segment traces into channel-id order: dst[i] = src[perm[i]] (rows of npts).
__global__ void __launch_bounds__(kStrThreads)
stream_gather_kernel(const float *__restrict__ src, const int *__restrict__ perm, int npts,
                     float *__restrict__ dst) {
  const int i = blockIdx.y;
  const float *a = src + (size_t)perm[i] * npts;
  float *b = dst + (size_t)i * npts;
  for (int t = blockIdx.x * kStrThreads + threadIdx.x; t < npts; t += gridDim.x * kStrThreads)
    b[t] = a[t];
}

// out[p][l] = (float)(acc[slot][l] / count); 0 for a pair never updated.
__global__ void __launch_bounds__(kStrThreads)
stream_linear_kernel(const double *__restrict__ acc, const int *__restrict__ slot,
                     const long long *__restrict__ count, long long npairs, int ncc,
                     float *__restrict__ out) {
  const int l = blockIdx.x * kStrThreads + threadIdx.x;
  if (l >= ncc) return;
  for (long long p = blockIdx.y; p < npairs; p += gridDim.y) {
    const long long c = count[p];
    out[p * ncc + l] = c > 0 ? (float)(acc[(long long)slot[p] * ncc + l] / (double)c) : 0.f;
  }
}

// PWS / TF-PWS prestack of each pair's last T records (ring order), float
// arithmetic of the reference host preparation: group sums in record order;
// PWS rows are means, lin = sum_g mean_g (n_g/T); TF-PWS rows are sums,
// lin = sum_g sum_g (1/T).
__global__ void __launch_bounds__(kStrThreads)
stream_prestack_kernel(const float *__restrict__ hist, int history, int ncc, long long npairs,
                       const int *__restrict__ slot, const int *__restrict__ start,
                       const int *__restrict__ nrec, const long long *__restrict__ goff, int B,
                       int tf, float *__restrict__ pre, float *__restrict__ lin) {
  const int l = blockIdx.x * kStrThreads + threadIdx.x;
  if (l >= ncc) return;
  for (long long p = blockIdx.y; p < npairs; p += gridDim.y) {
    const long long s = slot[p];
    const int t0 = start[p], n = nrec[p];
    const float inv_total = 1.0f / (float)n;
    float lsum = 0.f;
    long long g = goff[p];
    for (int r = 0; r < n; r += B, ++g) {
      const int cnt = min(B, n - r);
      float tr = 0.f;
      for (int i = 0; i < cnt; ++i) tr += hist[(s * history + (t0 + r + i) % history) * ncc + l];
      if (tf) {
        lsum += tr * inv_total;
      } else {
        tr *= 1.0f / (float)cnt;
        lsum += tr * ((float)cnt / (float)n);
      }
      pre[g * ncc + l] = tr;
    }
    lin[p * ncc + l] = lsum;
  }
}

static dim3 row_grid(int ncc, long long rows) {
  return dim3((unsigned)ceil_div((size_t)ncc, (size_t)kStrThreads),
              (unsigned)std::max<long long>(1, std::min<long long>(rows, 65535)));
}

}  // namespace fxc

using namespace fxc;

namespace {

using Clock = std::chrono::steady_clock;

struct StreamJob {
  long long ticket = 0;
  long long seq = -1;  // admission order of an admitted segment
  std::vector<int> chan;
  const float *h_traces = nullptr;
  Clock::time_point t_submit;
};

struct StreamResult {
  bool done = false;
  int status = FXC_OK;
  std::string msg;
  double latency_ms = 0.0;
};

struct StreamLane {
  hipStream_t st = nullptr;
  fxc_s2s_plan *s2s = nullptr;
  fxc_xc_plan *xc = nullptr;
  float *d_raw = nullptr;  // traces as submitted
  float *d_traces = nullptr, *d_spec = nullptr, *d_cc = nullptr;
  int *d_slot = nullptr, *d_hpos = nullptr, *d_perm = nullptr, *d_close = nullptr;
  int *h_slot = nullptr, *h_hpos = nullptr, *h_perm = nullptr, *h_close = nullptr;  // pinned
  hipEvent_t ev_cc = nullptr, ev_done = nullptr;
  std::thread th;
};

}  // namespace

struct fxc_stream {
  int dev = 0;
  fxc_s2s_config cfg{};
  std::vector<float> responses, freq_lows;
  int nchan = 0, half_cc = 0, ncc = 0, precision = 0, history = 0, nstep = 0, nspec = 0;
  int max_seg_chan = 0;
  long long max_pairs = 0, chunk_rows = 0;
  size_t lane_bytes = 0, state_bytes = 0;
  // pair state
  double *d_acc = nullptr;
  float *d_hist = nullptr;
  // running PWS (sub-stacks of pws_B records; 0 = off)
  int pws_B = 0;
  float2 *d_pws_S = nullptr;
  float *d_pws_open = nullptr;
  std::vector<int> open_cnt;  // records in each slot's open sub-stack
  std::unordered_map<long long, int> slot_of;
  std::vector<long long> count;  // records assigned per slot (submission order)
  hipStream_t upd = nullptr;
  std::mutex upd_mu;
  // queue
  std::mutex mu;
  std::condition_variable cv_q, cv_done, cv_seq;
  long long admitted = 0, next_update = 0;  // updates are applied in admission order
  std::deque<StreamJob> queue;
  std::map<long long, StreamResult> results;
  long long next_ticket = 1, submitted = 0, finished = 0;
  bool stop = false;
  // a segment that failed after its pairs were admitted leaves their counts
  // ahead of the accumulators: every later call reports it
  std::string broken;
  std::vector<StreamLane> lanes;
};

namespace {

int lane_run(fxc_stream *s, StreamLane &L, const StreamJob &job, const std::vector<int> &src,
             const std::vector<int> &rec, const std::vector<int> &slots,
             const std::vector<int> &hpos, const std::vector<int> &order,
             const std::vector<int> &close) {
  const int n = (int)job.chan.size();
  const long long np = (long long)src.size();
  const long long nclose = (long long)close.size() / 2;
  // traces in channel-id order, so every pair's lower id has the lower local
  // index and the pairs fill the upper-triangle station tiles only
  FXC_HIP(hipMemcpyAsync(L.d_raw, job.h_traces, (size_t)n * s->cfg.npts * sizeof(float),
                         hipMemcpyHostToDevice, L.st));
  std::memcpy(L.h_perm, order.data(), (size_t)n * sizeof(int));
  FXC_HIP(hipMemcpyAsync(L.d_perm, L.h_perm, (size_t)n * sizeof(int), hipMemcpyHostToDevice, L.st));
  hipLaunchKernelGGL(stream_gather_kernel,
                     dim3((unsigned)std::min<size_t>(64, ceil_div((size_t)s->cfg.npts, kStrThreads)),
                          (unsigned)n),
                     dim3(kStrThreads), 0, L.st, L.d_raw, L.d_perm, s->cfg.npts, L.d_traces);
  FXC_HIP(hipGetLastError());
  std::vector<int> steps((size_t)s->nstep);
  for (int i = 0; i < s->nstep; ++i) steps[i] = i;
  FXC_TRY(fxc_s2s_process(L.s2s, L.d_traces, n, steps.data(), s->nstep, L.d_spec, L.st));
  FXC_TRY(xc_set_pairs_on(L.xc, n, src.data(), rec.data(), np, L.st));
  FXC_TRY(fxc_xc_correlate(L.xc, L.d_spec, L.d_cc, L.st));
  std::memcpy(L.h_slot, slots.data(), (size_t)np * sizeof(int));
  std::memcpy(L.h_hpos, hpos.data(), (size_t)np * sizeof(int));
  FXC_HIP(hipMemcpyAsync(L.d_slot, L.h_slot, (size_t)np * sizeof(int), hipMemcpyHostToDevice, L.st));
  FXC_HIP(hipMemcpyAsync(L.d_hpos, L.h_hpos, (size_t)np * sizeof(int), hipMemcpyHostToDevice, L.st));
  if (nclose > 0) {
    std::memcpy(L.h_close, close.data(), (size_t)nclose * 2 * sizeof(int));
    FXC_HIP(hipMemcpyAsync(L.d_close, L.h_close, (size_t)nclose * 2 * sizeof(int),
                           hipMemcpyHostToDevice, L.st));
  }
  FXC_HIP(hipEventRecord(L.ev_cc, L.st));
  // State updates are enqueued in admission order: a history position is
  // count mod history, so a record may only land after every earlier record
  // of its pair (a slow lane could otherwise still hold record r when record
  // r + history lands), and the float64 sums add in the order of a
  // sequential run.  Only this small update waits; the transforms above
  // overlap freely.
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_seq.wait(lk, [&] { return s->next_update == job.seq; });
  }
  int rc = FXC_OK;
  {
    std::lock_guard<std::mutex> g(s->upd_mu);
    if (hipStreamWaitEvent(s->upd, L.ev_cc, 0) != hipSuccess) {
      set_error("stream: hipStreamWaitEvent failed");
      rc = FXC_ERR_HIP;
    } else {
      hipLaunchKernelGGL(stream_update_kernel, row_grid(s->ncc, np), dim3(kStrThreads), 0,
                         s->upd, L.d_cc, s->ncc, np, L.d_slot, L.d_hpos, s->d_acc, s->d_hist,
                         s->history, s->d_pws_open);
      if (hipGetLastError() != hipSuccess) {
        set_error("stream: update launch failed");
        rc = FXC_ERR_HIP;
      }
      // sub-stacks this segment completes: [slots | counts] in d_close; with
      // sub-stacks of one record every pair closes, from its new correlation
      if (rc == FXC_OK && s->pws_B == 1)
        rc = stream_pws_close(s->d_pws_S, nullptr, L.d_slot, nullptr, np, s->ncc, s->upd, L.d_cc);
      else if (rc == FXC_OK && nclose > 0)
        rc = stream_pws_close(s->d_pws_S, s->d_pws_open, L.d_close, L.d_close + nclose, nclose,
                              s->ncc, s->upd);
      if (rc == FXC_OK && hipEventRecord(L.ev_done, s->upd) != hipSuccess) {
        set_error("stream: update event failed");
        rc = FXC_ERR_HIP;
      }
    }
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->next_update;
  }
  s->cv_seq.notify_all();
  FXC_TRY(rc);
  FXC_HIP(hipEventSynchronize(L.ev_done));
  return FXC_OK;
}

void lane_main(fxc_stream *s, StreamLane *L) {
  (void)hipSetDevice(s->dev);
  for (;;) {
    StreamJob job;
    std::vector<int> src, rec, slots, hpos, order, close;
    int status = FXC_OK;
    std::string msg;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_q.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->queue.empty()) return;  // stopping, queue drained
      job = std::move(s->queue.front());
      s->queue.pop_front();
      // admission, in submission order: pairs (lower channel id first), slots
      const int n = (int)job.chan.size();
      order.resize((size_t)n);
      for (int i = 0; i < n; ++i) order[i] = i;
      std::sort(order.begin(), order.end(),
                [&](int a, int b) { return job.chan[a] < job.chan[b]; });
      const long long np = (long long)n * (n - 1) / 2;
      src.reserve((size_t)np);
      rec.reserve((size_t)np);
      std::vector<long long> keys;
      keys.reserve((size_t)np);
      long long fresh = 0;
      // local indices are positions in channel-id order (the lane gathers the
      // traces into that order)
      for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b) {
          src.push_back(a);
          rec.push_back(b);
          const long long key = (long long)job.chan[order[a]] * s->nchan + job.chan[order[b]];
          keys.push_back(key);
          if (!s->slot_of.count(key)) ++fresh;
        }
      if ((long long)s->slot_of.size() + fresh > s->max_pairs) {
        status = FXC_ERR_ALLOC;
        msg = "stream: pair state full (" + std::to_string(s->slot_of.size()) + " + " +
              std::to_string(fresh) + " new > " + std::to_string(s->max_pairs) + " slots)";
      } else {
        slots.resize((size_t)np);
        hpos.resize((size_t)np);
        for (long long p = 0; p < np; ++p) {
          auto it = s->slot_of.find(keys[p]);
          int sl;
          if (it == s->slot_of.end()) {
            sl = (int)s->slot_of.size();
            s->slot_of.emplace(keys[p], sl);
            s->count.push_back(0);
            s->open_cnt.push_back(0);
          } else {
            sl = it->second;
          }
          slots[p] = sl;
          hpos[p] = s->history > 0 ? (int)(s->count[sl] % s->history) : 0;
          ++s->count[sl];
          if (s->pws_B > 0 && ++s->open_cnt[sl] == s->pws_B) {
            close.push_back(sl);
            s->open_cnt[sl] = 0;
          }
        }
        // [slots | counts] of the sub-stacks this segment closes
        const size_t nc = close.size();
        close.resize(2 * nc, s->pws_B);
        job.seq = s->admitted++;
      }
    }
    if (status == FXC_OK) {
      status = lane_run(s, *L, job, src, rec, slots, hpos, order, close);
      if (status != FXC_OK) {
        // a lane that failed before its update turn still passes the turn on
        {
          std::unique_lock<std::mutex> lk(s->mu);
          s->cv_seq.wait(lk, [&] { return s->next_update >= job.seq; });
          if (s->next_update == job.seq) ++s->next_update;
        }
        s->cv_seq.notify_all();
        msg = last_error();
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->broken.empty()) s->broken = "stream: segment " + std::to_string(job.ticket) +
                                           " failed after admission: " + msg;
      }
    }
    const double ms =
        std::chrono::duration<double, std::milli>(Clock::now() - job.t_submit).count();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      StreamResult &r = s->results[job.ticket];
      r.done = true;
      r.status = status;
      r.msg = msg;
      r.latency_ms = ms;
      ++s->finished;
    }
    s->cv_done.notify_all();
  }
}

void stream_free(fxc_stream *s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_q.notify_all();
  for (auto &L : s->lanes)
    if (L.th.joinable()) L.th.join();
  for (auto &L : s->lanes) {
    if (L.s2s) fxc_s2s_plan_destroy(L.s2s);
    if (L.xc) fxc_xc_plan_destroy(L.xc);
    dfree(L.d_raw);
    dfree(L.d_traces);
    dfree(L.d_spec);
    dfree(L.d_cc);
    dfree(L.d_slot);
    dfree(L.d_hpos);
    dfree(L.d_perm);
    dfree(L.d_close);
    for (int *h : {L.h_slot, L.h_hpos, L.h_perm, L.h_close})
      if (h) (void)hipHostFree(h);
    if (L.ev_cc) (void)hipEventDestroy(L.ev_cc);
    if (L.ev_done) (void)hipEventDestroy(L.ev_done);
    if (L.st) (void)hipStreamDestroy(L.st);
  }
  dfree(s->d_acc);
  dfree(s->d_hist);
  dfree(s->d_pws_S);
  dfree(s->d_pws_open);
  if (s->upd) (void)hipStreamDestroy(s->upd);
  delete s;
}

// Lane set-up: plans, buffers, and one pass over a silent segment of the
// largest size so every workspace is allocated before the first request.
int lane_init(fxc_stream *s, StreamLane &L) {
  FXC_HIP(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
  FXC_HIP(hipEventCreateWithFlags(&L.ev_cc, hipEventDisableTiming));
  FXC_HIP(hipEventCreateWithFlags(&L.ev_done, hipEventDisableTiming));
  FXC_TRY(fxc_s2s_plan_create(&L.s2s, &s->cfg));
  FXC_TRY(fxc_xc_plan_create(&L.xc, s->nspec, s->nstep, s->cfg.delta, s->half_cc, s->precision));
  FXC_TRY(fxc_xc_set_chunk_rows(L.xc, s->chunk_rows));
  const int n = s->max_seg_chan;
  const long long np = (long long)n * (n - 1) / 2;
  FXC_TRY(dmalloc((void **)&L.d_raw, (size_t)n * s->cfg.npts * sizeof(float)));
  FXC_TRY(dmalloc((void **)&L.d_traces, (size_t)n * s->cfg.npts * sizeof(float)));
  FXC_TRY(dmalloc((void **)&L.d_perm, (size_t)n * sizeof(int)));
  FXC_TRY(dmalloc((void **)&L.d_close, (size_t)np * 2 * sizeof(int)));
  FXC_HIP(hipHostMalloc((void **)&L.h_perm, (size_t)n * sizeof(int), hipHostMallocDefault));
  FXC_HIP(hipHostMalloc((void **)&L.h_close, (size_t)np * 2 * sizeof(int), hipHostMallocDefault));
  FXC_TRY(dmalloc((void **)&L.d_spec, (size_t)s->nstep * n * s->nspec * 2 * sizeof(float)));
  FXC_TRY(dmalloc((void **)&L.d_cc, (size_t)np * s->ncc * sizeof(float)));
  FXC_TRY(dmalloc((void **)&L.d_slot, (size_t)np * sizeof(int)));
  FXC_TRY(dmalloc((void **)&L.d_hpos, (size_t)np * sizeof(int)));
  FXC_HIP(hipHostMalloc((void **)&L.h_slot, (size_t)np * sizeof(int), hipHostMallocDefault));
  FXC_HIP(hipHostMalloc((void **)&L.h_hpos, (size_t)np * sizeof(int), hipHostMallocDefault));
  // warm-up: zero traces through sac2spec and xc (no state update)
  FXC_HIP(hipMemsetAsync(L.d_traces, 0, (size_t)n * s->cfg.npts * sizeof(float), L.st));
  std::vector<int> steps((size_t)s->nstep), src, rec;
  for (int i = 0; i < s->nstep; ++i) steps[i] = i;
  for (int a = 0; a < n; ++a)
    for (int b = a + 1; b < n; ++b) {
      src.push_back(a);
      rec.push_back(b);
    }
  FXC_TRY(fxc_s2s_process(L.s2s, L.d_traces, n, steps.data(), s->nstep, L.d_spec, L.st));
  FXC_TRY(xc_set_pairs_on(L.xc, n, src.data(), rec.data(), np, L.st));
  FXC_TRY(fxc_xc_correlate(L.xc, L.d_spec, L.d_cc, L.st));
  FXC_HIP(hipStreamSynchronize(L.st));
  return FXC_OK;
}

}  // namespace

extern "C" int fxc_stream_create(fxc_stream **out, const fxc_s2s_config *cfg, int nchan,
                                 int half_cc, int precision, int history, long long max_pairs,
                                 int max_seg_chan, int lanes) {
  FXC_CHECK_ARG(out && cfg && nchan >= 2 && half_cc >= 0 && history >= 0 && max_pairs >= 1 &&
                    max_seg_chan >= 2 && lanes >= 1 && lanes <= 16,
                "fxc_stream_create: invalid arguments");
  FXC_CHECK_ARG(cfg->npts >= cfg->seg && cfg->seg > 0 && cfg->shift > 0 && cfg->num_ch >= 1 &&
                    cfg->filter_count >= 1 && cfg->responses && cfg->freq_lows,
                "fxc_stream_create: invalid sac2spec configuration");
  FXC_CHECK_ARG(max_seg_chan % cfg->num_ch == 0,
                "fxc_stream_create: max_seg_chan must be a multiple of num_ch");
  // lanes in flight write distinct history rows of a pair only if the ring
  // is at least as long as the number of lanes
  FXC_CHECK_ARG(history == 0 || history >= lanes,
                "fxc_stream_create: history must be 0 or at least the lane count");
  fxc_stream *s = new fxc_stream();
  FXC_HIP(hipGetDevice(&s->dev));
  s->cfg = *cfg;
  const size_t rlen = (size_t)cfg->filter_count * (cfg->nfft_filter / 2 + 1);
  s->responses.assign(cfg->responses, cfg->responses + rlen);
  s->freq_lows.assign(cfg->freq_lows, cfg->freq_lows + cfg->filter_count);
  s->cfg.responses = s->responses.data();
  s->cfg.freq_lows = s->freq_lows.data();
  s->nchan = nchan;
  s->nstep = (cfg->npts - cfg->seg) / cfg->shift + 1;
  s->nspec = cfg->nfft_out / 2 + 1;
  // 2H+1 must fit nfft = 2(nspec-1): H <= nspec - 2
  s->half_cc = std::min(half_cc, s->nspec - 2);
  s->ncc = 2 * s->half_cc + 1;
  s->precision = precision;
  s->history = history;
  s->max_pairs = max_pairs;
  s->max_seg_chan = max_seg_chan;
  const long long np = (long long)max_seg_chan * (max_seg_chan - 1) / 2;
  s->chunk_rows = std::min<long long>(np, 65536);
  s->state_bytes = (size_t)max_pairs * s->ncc * (sizeof(double) + (size_t)history * sizeof(float));
  int rc = FXC_OK;
  auto fail = [&](int code) {
    stream_free(s);
    return code;
  };
  if (hipStreamCreateWithFlags(&s->upd, hipStreamNonBlocking) != hipSuccess)
    return fail((set_error("fxc_stream_create: stream creation failed"), FXC_ERR_HIP));
  if ((rc = dmalloc((void **)&s->d_acc, (size_t)max_pairs * s->ncc * sizeof(double))) != FXC_OK)
    return fail(rc);
  if (hipMemset(s->d_acc, 0, (size_t)max_pairs * s->ncc * sizeof(double)) != hipSuccess)
    return fail((set_error("fxc_stream_create: clearing the state failed"), FXC_ERR_HIP));
  if (history > 0 &&
      (rc = dmalloc((void **)&s->d_hist, (size_t)max_pairs * history * s->ncc * sizeof(float))) !=
          FXC_OK)
    return fail(rc);
  s->lanes.resize((size_t)lanes);
  size_t before = 0, after = 0, total = 0;
  (void)hipMemGetInfo(&before, &total);
  for (auto &L : s->lanes)
    if ((rc = lane_init(s, L)) != FXC_OK) return fail(rc);
  (void)hipMemGetInfo(&after, &total);
  s->lane_bytes = before > after ? (before - after) / (size_t)lanes : 0;
  for (auto &L : s->lanes) L.th = std::thread(lane_main, s, &L);
  *out = s;
  return FXC_OK;
}

extern "C" int fxc_stream_submit(fxc_stream *s, const int *h_chan, int nchan_seg,
                                 const float *h_traces, long long *ticket) {
  FXC_CHECK_ARG(s && h_chan && h_traces && ticket && nchan_seg >= 2,
                "fxc_stream_submit: invalid arguments");
  FXC_CHECK_ARG(nchan_seg <= s->max_seg_chan, "fxc_stream_submit: %d channels exceed the lanes' %d",
                nchan_seg, s->max_seg_chan);
  FXC_CHECK_ARG(nchan_seg % s->cfg.num_ch == 0,
                "fxc_stream_submit: channel count must be a multiple of num_ch");
  StreamJob job;
  job.chan.assign(h_chan, h_chan + nchan_seg);
  std::vector<int> sorted = job.chan;
  std::sort(sorted.begin(), sorted.end());
  for (int i = 0; i < nchan_seg; ++i) {
    FXC_CHECK_ARG(sorted[i] >= 0 && sorted[i] < s->nchan, "fxc_stream_submit: channel id out of range");
    FXC_CHECK_ARG(i == 0 || sorted[i] != sorted[i - 1], "fxc_stream_submit: duplicate channel %d",
                  sorted[i]);
  }
  job.h_traces = h_traces;
  job.t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    FXC_CHECK_ARG(s->broken.empty(), "%s", s->broken.c_str());
    job.ticket = s->next_ticket++;
    *ticket = job.ticket;
    s->results[job.ticket] = StreamResult();
    s->queue.push_back(std::move(job));
    ++s->submitted;
  }
  s->cv_q.notify_one();
  return FXC_OK;
}

extern "C" int fxc_stream_wait(fxc_stream *s, long long ticket, double *latency_ms) {
  FXC_CHECK_ARG(s, "fxc_stream_wait: null stream");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->results.find(ticket);
  FXC_CHECK_ARG(it != s->results.end(), "fxc_stream_wait: unknown ticket %lld", ticket);
  s->cv_done.wait(lk, [&] { return s->results[ticket].done; });
  StreamResult r = s->results[ticket];
  s->results.erase(ticket);
  lk.unlock();
  if (latency_ms) *latency_ms = r.latency_ms;
  if (r.status != FXC_OK) {
    set_error("%s", r.msg.c_str());
    return r.status;
  }
  return FXC_OK;
}

extern "C" int fxc_stream_info(fxc_stream *s, long long *pairs_used, long long *segments_done,
                               long long *state_bytes, long long *lane_bytes) {
  FXC_CHECK_ARG(s, "fxc_stream_info: null stream");
  std::lock_guard<std::mutex> lk(s->mu);
  if (pairs_used) *pairs_used = (long long)s->slot_of.size();
  if (segments_done) *segments_done = s->finished;
  if (state_bytes) *state_bytes = (long long)s->state_bytes;
  if (lane_bytes) *lane_bytes = (long long)s->lane_bytes;
  return FXC_OK;
}

extern "C" int fxc_stream_set_pws(fxc_stream *s, int substack) {
  FXC_CHECK_ARG(s && substack >= 1, "fxc_stream_set_pws: invalid arguments");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    FXC_CHECK_ARG(s->submitted == 0 && s->pws_B == 0,
                  "fxc_stream_set_pws: must be called once, before the first segment");
  }
  // sub-stacks of one record close as they arrive: no open sums are kept
  const size_t n = (size_t)s->max_pairs * s->ncc;
  FXC_TRY(dmalloc((void **)&s->d_pws_S, n * sizeof(float2)));
  FXC_HIP(hipMemset(s->d_pws_S, 0, n * sizeof(float2)));
  s->state_bytes += n * sizeof(float2);
  if (substack > 1) {
    FXC_TRY(dmalloc((void **)&s->d_pws_open, n * sizeof(float)));
    FXC_HIP(hipMemset(s->d_pws_open, 0, n * sizeof(float)));
    s->state_bytes += n * sizeof(float);
  }
  s->pws_B = substack;
  return FXC_OK;
}

static int stream_read_impl(fxc_stream *s, int kind, const int *h_src, const int *h_rec,
                            long long npairs, int substack, int band_limited, float fmin,
                            float fmax, float taper_hz, const fxc_stack_spec *spec, float *d_out,
                            long long *h_count, void *stream);

extern "C" int fxc_stream_read(fxc_stream *s, int kind, const int *h_src, const int *h_rec,
                               long long npairs, int substack, int band_limited, float fmin,
                               float fmax, float taper_hz, float *d_out, long long *h_count,
                               void *stream) {
  return stream_read_impl(s, kind, h_src, h_rec, npairs, substack, band_limited, fmin, fmax,
                          taper_hz, nullptr, d_out, h_count, stream);
}

extern "C" int fxc_stream_read_spec(fxc_stream *s, int kind, const int *h_src, const int *h_rec,
                                    long long npairs, int substack, int band_limited, float fmin,
                                    float fmax, float taper_hz, const fxc_stack_spec *spec,
                                    float *d_out, long long *h_count, void *stream) {
  return stream_read_impl(s, kind, h_src, h_rec, npairs, substack, band_limited, fmin, fmax,
                          taper_hz, spec, d_out, h_count, stream);
}

static int stream_read_impl(fxc_stream *s, int kind, const int *h_src, const int *h_rec,
                            long long npairs, int substack, int band_limited, float fmin,
                            float fmax, float taper_hz, const fxc_stack_spec *spec, float *d_out,
                            long long *h_count, void *stream) {
  FXC_CHECK_ARG(s && h_src && h_rec && d_out && npairs >= 1 && kind >= 0 && kind <= 2,
                "fxc_stream_read: invalid arguments");
  FXC_CHECK_ARG(!spec || (std::isfinite(spec->power) && spec->power > 0.f &&
                          (spec->nvoices == 0 || spec->nvoices >= 2)),
                "fxc_stream_read: invalid spec");
  const float power = spec ? spec->power : 1.f;
  // PWS from the running state when the engine keeps one, else (and for
  // TF-PWS) from the history ring
  const bool running_pws = kind == 1 && s->pws_B > 0;
  FXC_CHECK_ARG(!running_pws || substack <= 1 || substack == s->pws_B,
                "fxc_stream_read: the running PWS state has sub-stacks of %d records", s->pws_B);
  FXC_CHECK_ARG(kind == 0 || running_pws || s->history > 0,
                "fxc_stream_read: PWS / TF-PWS need a history or the running PWS state");
  hipStream_t st = (hipStream_t)stream;
  // every segment submitted so far is applied first
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->finished == s->submitted; });
  FXC_CHECK_ARG(s->broken.empty(), "%s", s->broken.c_str());
  std::vector<int> slot((size_t)npairs, 0), start((size_t)npairs, 0), nrec((size_t)npairs, 0);
  std::vector<long long> cnt((size_t)npairs, 0);
  std::vector<int> nopen((size_t)npairs, 0);
  for (long long p = 0; p < npairs; ++p) {
    FXC_CHECK_ARG(h_src[p] >= 0 && h_rec[p] < s->nchan && h_src[p] < h_rec[p],
                  "fxc_stream_read: pair %lld must be channel ids src < rec", p);
    auto it = s->slot_of.find((long long)h_src[p] * s->nchan + h_rec[p]);
    if (it != s->slot_of.end()) {
      slot[p] = it->second;
      cnt[p] = s->count[it->second];
      nopen[p] = s->open_cnt[it->second];
    }
    if (running_pws) {
      FXC_CHECK_ARG(cnt[p] > 0, "fxc_stream_read: pair %lld has no records", p);
    } else if (kind != 0) {
      FXC_CHECK_ARG(cnt[p] > 0, "fxc_stream_read: pair %lld has no records", p);
      nrec[p] = (int)std::min<long long>(cnt[p], s->history);
      start[p] = (int)((cnt[p] - nrec[p]) % s->history);
    }
  }
  lk.unlock();
  if (h_count) std::memcpy(h_count, cnt.data(), (size_t)npairs * sizeof(long long));
  FXC_HIP(hipStreamSynchronize(s->upd));
  struct Tmp {
    std::vector<void *> v;
    ~Tmp() {
      for (void *q : v) dfree(q);
    }
  } tmp;
  auto dev = [&](void **d, size_t bytes) -> int {
    FXC_TRY(dmalloc(d, std::max<size_t>(bytes, 1)));
    tmp.v.push_back(*d);
    return FXC_OK;
  };
  int *d_slot = nullptr;
  FXC_TRY(dev((void **)&d_slot, (size_t)npairs * sizeof(int)));
  FXC_HIP(hipMemcpyAsync(d_slot, slot.data(), (size_t)npairs * sizeof(int), hipMemcpyHostToDevice, st));
  if (running_pws) {
    long long *d_cnt = nullptr;
    int *d_nopen = nullptr;
    FXC_TRY(dev((void **)&d_cnt, (size_t)npairs * sizeof(long long)));
    FXC_TRY(dev((void **)&d_nopen, (size_t)npairs * sizeof(int)));
    FXC_HIP(hipMemcpyAsync(d_cnt, cnt.data(), (size_t)npairs * sizeof(long long),
                           hipMemcpyHostToDevice, st));
    FXC_HIP(hipMemcpyAsync(d_nopen, nopen.data(), (size_t)npairs * sizeof(int),
                           hipMemcpyHostToDevice, st));
    FXC_TRY(stream_pws_read(s->d_pws_S, s->d_pws_open, s->d_acc, d_slot, d_cnt, d_nopen, npairs,
                            s->ncc, power, d_out, st));
    FXC_HIP(hipStreamSynchronize(st));
    return FXC_OK;
  }
  if (kind == 0) {
    long long *d_cnt = nullptr;
    FXC_TRY(dev((void **)&d_cnt, (size_t)npairs * sizeof(long long)));
    FXC_HIP(hipMemcpyAsync(d_cnt, cnt.data(), (size_t)npairs * sizeof(long long),
                           hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(stream_linear_kernel, row_grid(s->ncc, npairs), dim3(kStrThreads), 0, st,
                       s->d_acc, d_slot, d_cnt, npairs, s->ncc, d_out);
    FXC_HIP(hipGetLastError());
    FXC_HIP(hipStreamSynchronize(st));
    return FXC_OK;
  }
  const int B = substack < 2 ? 1 : substack;
  std::vector<long long> goff((size_t)npairs + 1, 0);
  std::vector<float> w;
  for (long long p = 0; p < npairs; ++p) {
    const int n = nrec[p];
    const float inv_total = 1.0f / (float)n;
    for (int r = 0; r < n; r += B) {
      const int c = std::min(B, n - r);
      w.push_back(kind == 2 ? (float)c * inv_total : (float)c / (float)n);
    }
    goff[p + 1] = (long long)w.size();
  }
  int *d_start = nullptr, *d_nrec = nullptr;
  long long *d_goff = nullptr;
  float *d_pre = nullptr, *d_lin = nullptr, *d_w = nullptr;
  FXC_TRY(dev((void **)&d_start, (size_t)npairs * sizeof(int)));
  FXC_TRY(dev((void **)&d_nrec, (size_t)npairs * sizeof(int)));
  FXC_TRY(dev((void **)&d_goff, (size_t)(npairs + 1) * sizeof(long long)));
  FXC_TRY(dev((void **)&d_pre, (size_t)w.size() * s->ncc * sizeof(float)));
  FXC_TRY(dev((void **)&d_lin, (size_t)npairs * s->ncc * sizeof(float)));
  FXC_TRY(dev((void **)&d_w, w.size() * sizeof(float)));
  FXC_HIP(hipMemcpyAsync(d_start, start.data(), (size_t)npairs * sizeof(int), hipMemcpyHostToDevice, st));
  FXC_HIP(hipMemcpyAsync(d_nrec, nrec.data(), (size_t)npairs * sizeof(int), hipMemcpyHostToDevice, st));
  FXC_HIP(hipMemcpyAsync(d_goff, goff.data(), (size_t)(npairs + 1) * sizeof(long long),
                         hipMemcpyHostToDevice, st));
  FXC_HIP(hipMemcpyAsync(d_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(stream_prestack_kernel, row_grid(s->ncc, npairs), dim3(kStrThreads), 0, st,
                     s->d_hist, s->history, s->ncc, npairs, d_slot, d_start, d_nrec, d_goff, B,
                     kind == 2 ? 1 : 0, d_pre, d_lin);
  FXC_HIP(hipGetLastError());
  if (kind == 1)
    FXC_TRY(fxc_pws_spec(d_pre, d_w, goff.data(), d_lin, (int)npairs, s->ncc, spec, d_out, st));
  else
    FXC_TRY(fxc_tfpws_spec(d_pre, d_w, goff.data(), d_lin, (int)npairs, s->ncc, band_limited,
                           s->cfg.delta, fmin, fmax, taper_hz, spec, d_out, st));
  FXC_HIP(hipStreamSynchronize(st));
  return FXC_OK;
}

extern "C" int fxc_stream_destroy(fxc_stream *s) {
  if (!s) return FXC_OK;
  stream_free(s);
  return FXC_OK;
}
