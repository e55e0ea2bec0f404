// Station-pair planning: every candidate pair's geometry and filters.
//
// Replaces build_valid_paths / _build_path_record_if_valid /
// azimuth_in_ranges / distkm_az_baz_rudoe
// (fastxc/inventory/planner.py:294-312, 564-617, 833-1000), which the
// reference runs as a Python loop over all N(N+1)/2 node pairs -- the part of
// inventory that grows with the square of the station count.
//
// Host code: the geometry is float64 libm arithmetic in the reference's
// operation order, so every value equals the reference's bit for bit (the
// same C library computes atan, atan2, sin, cos, tan, sqrt and the squares,
// which Python takes with pow, and nothing is contracted into fused
// multiply-adds).  Per-node terms (direction
// cosines, latitude trigonometry) are computed once; source rows are handed
// to host threads from a shared counter and their pairs concatenated in row
// order, which is the reference's (src, rec) order.
#pragma clang fp contract(off)

#include "fxc_common.hpp"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace fxc {
namespace {

constexpr double kAkm = 6378.137;
constexpr double kF = 1.0 / 298.257223563;
constexpr double kFltMax = 3.4028234663852886e38;

struct Node {
  bool geo = false;
  double lat = 0, lon = 0;   // degrees as given
  double lat_rad = 0, lon_rad = 0;
  double v[8];               // direction-cosine terms (a b c d e f g h)
  double cos_l = 0, sin_l = 0, tan_l = 0;  // latitude trigonometry (poles fixed)
};

// Python's x ** 2: libm pow(|x|, 2.0), which is not always x * x (it can be
// one ulp off); the exponent is read through a volatile so the compiler does
// not turn the call into a multiplication.
double sq(double x) {
  volatile double two = 2.0;
  return std::pow(std::fabs(x), two);
}

void node_terms(Node &n, double d2r, double one_minus_ec2) {
  n.lat_rad = d2r * (n.lat != 0.0 ? n.lat : 1.0e-10);
  n.lon_rad = d2r * n.lon;
  const double geoc = (n.lat == 90.0 || n.lat == -90.0)
                          ? n.lat * d2r
                          : std::atan(one_minus_ec2 * std::tan(n.lat_rad));
  const double d = std::sin(n.lon_rad), e = -std::cos(n.lon_rad);
  const double f = -std::cos(geoc), c = std::sin(geoc);
  n.v[0] = f * e;
  n.v[1] = -f * d;
  n.v[2] = c;
  n.v[3] = d;
  n.v[4] = e;
  n.v[5] = f;
  n.v[6] = -c * e;
  n.v[7] = c * d;
  if (n.lat == 90.0) {
    n.cos_l = 0.0, n.sin_l = 1.0, n.tan_l = kFltMax;
  } else if (n.lat == -90.0) {
    n.cos_l = 0.0, n.sin_l = -1.0, n.tan_l = -kFltMax;
  } else {
    n.cos_l = std::cos(n.lat_rad);
    n.sin_l = std::sin(n.lat_rad);
    n.tan_l = n.sin_l / n.cos_l;
  }
}

double azimuth(const Node &p, const Node &q, double r2d) {
  const double ss = sq(q.v[0] - p.v[3]) + sq(q.v[1] - p.v[4]) + sq(q.v[2]) - 2.0;
  const double sc = sq(q.v[0] - p.v[6]) + sq(q.v[1] - p.v[7]) + sq(q.v[2] - p.v[5]) - 2.0;
  double az = std::atan2(ss, sc) * r2d;
  if (az < 0.0 && std::fabs(az) < 1.0e-8) az = 0.0;
  if (az < 0.0) az += 360.0;
  return az;
}

struct Consts {
  double ec2, one_minus_ec2, d2r, r2d, e_prime, pi;
};

double rudoe_km(const Node &k, const Node &i, const Consts &C) {
  const double cos_k = k.cos_l, sin_k = k.sin_l, tan_k = k.tan_l;
  const double cos_i = i.cos_l, sin_i = i.sin_l, tan_i = i.tan_l;
  const double one_e = 1.0 + C.e_prime;
  const double al = tan_i / (one_e * tan_k) +
                    C.ec2 * std::sqrt((one_e + sq(tan_i)) / (one_e + sq(tan_k)));
  const double dl = i.lon_rad - k.lon_rad;
  const double a12 = std::atan2(std::sin(dl), (al - std::cos(dl)) * sin_k);
  const double cos_a12 = std::cos(a12), sin_a12 = std::sin(a12);
  const double e1 = C.e_prime * (sq(cos_k * cos_a12) + sq(sin_k));
  const double e2 = e1 * e1;
  const double e3 = e1 * e2;
  const double c0 = 1.0 + 0.25 * e1 - (3.0 / 64.0) * e2 + (5.0 / 256.0) * e3;
  // the reference's 13/32 (the textbook series has 1/32)
  const double c2 = -(1.0 / 8.0) * e1 + (13.0 / 32.0) * e2 - (15.0 / 1024.0) * e3;
  const double c4 = -(1.0 / 256.0) * e2 + (3.0 / 1024.0) * e3;
  const double nu_k = kAkm / std::sqrt(1.0 - C.ec2 * sq(sin_k));
  const double nu_i = kAkm / std::sqrt(1.0 - C.ec2 * sq(sin_i));
  const double z_k = nu_k * C.one_minus_ec2 * sin_k;
  const double z_i = nu_i * C.one_minus_ec2 * sin_i;
  const double x_i = nu_i * cos_i * std::cos(dl);
  const double y_i = nu_i * cos_i * std::sin(dl);
  const double e1p = e1 + 1.0;
  const double root = std::sqrt(e1p);
  const double u1 = std::atan2(tan_k, root * cos_a12);
  const double u2 = std::atan2(nu_k * sin_k + e1p * (z_i - z_k),
                               root * (x_i * cos_a12 - y_i * sin_k * sin_a12));
  const double b0 = nu_k * std::sqrt(1.0 + C.e_prime * sq(cos_k * cos_a12)) / e1p;
  double du = u2 - u1;
  if (std::fabs(du) > C.pi)
    du = (du > 0.0) ? (C.pi * 2.0 - std::fabs(du)) : (std::fabs(du) - C.pi * 2.0);
  const double series =
      b0 * (c2 * (std::sin(2.0 * u2) - std::sin(2.0 * u1)) +
            c4 * (std::sin(4.0 * u2) - std::sin(4.0 * u1)));
  return std::fabs(b0 * c0 * du + series);
}

// gcarc, az, baz, dist of event ev and station st
void geometry(const Node &ev, const Node &st, const Consts &C, double out[4]) {
  if (ev.lon == st.lon && ev.lat == st.lat) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    return;
  }
  const double sc = ev.v[0] * st.v[0] + ev.v[1] * st.v[1] + ev.v[2] * st.v[2];
  const double sd = 0.5 * std::sqrt((sq(ev.v[0] - st.v[0]) + sq(ev.v[1] - st.v[1]) +
                                     sq(ev.v[2] - st.v[2])) *
                                    (sq(ev.v[0] + st.v[0]) + sq(ev.v[1] + st.v[1]) +
                                     sq(ev.v[2] + st.v[2])));
  double gcarc = std::atan2(sd, sc) * C.r2d;
  if (gcarc < 0.0) gcarc += 360.0;
  out[0] = gcarc;
  out[1] = azimuth(ev, st, C.r2d);
  out[2] = azimuth(st, ev, C.r2d);
  out[3] = st.lat_rad > 0.0 ? rudoe_km(ev, st, C) : rudoe_km(st, ev, C);
}

// Python's float x % y for y > 0
double pymod(double x, double y) {
  double r = std::fmod(x, y);
  if (r != 0.0) {
    if ((r < 0.0) != (y < 0.0)) r += y;
  } else {
    r = std::copysign(0.0, y);
  }
  return r;
}

bool in_ranges(double az, const double *ranges, int nranges) {
  az = pymod(az, 360.0);
  for (int r = 0; r < nranges; ++r) {
    const double lo = ranges[2 * r], hi = ranges[2 * r + 1];
    if (lo == 0.0 && hi == 360.0) return true;
    const double a = pymod(lo, 360.0), b = pymod(hi, 360.0);
    if (a <= b) {
      if (a <= az && az <= b) return true;
    } else if (az >= a || az <= b) {
      return true;
    }
  }
  return false;
}

struct PlanConfig {
  int allow_autocorr, autocorr_only, group_mode;  // 0 all, 1 same group, 2 cross group
  int use_distance;
  double min_km, max_km;
  int use_azimuth;
  std::vector<double> ranges;
};

struct Path {
  int src, rec;  // node indices
  unsigned char has_geo;
  double g[4];   // gcarc, az, baz, dist
};

}  // namespace

struct PathSet {
  std::vector<Path> paths;
  long long checks = 0;
};

}  // namespace fxc

using namespace fxc;

extern "C" int fxc_plan_paths(int nnodes, const double *lat, const double *lon,
                              const unsigned char *has_geo, const int *group, int allow_autocorr,
                              int autocorr_only, int group_mode, int use_distance, double min_km,
                              double max_km, int use_azimuth, const double *az_ranges,
                              int nranges, int nthreads, void **out_set, long long *npaths,
                              long long *nchecks) {
  FXC_CHECK_ARG(out_set && npaths && nchecks && nnodes >= 0, "fxc_plan_paths: invalid arguments");
  FXC_CHECK_ARG(nnodes == 0 || (lat && lon && has_geo && group),
                "fxc_plan_paths: node arrays required");
  FXC_CHECK_ARG(nnodes <= 9999, "fxc_plan_paths: at most 9999 nodes (4-digit ids)");
  FXC_CHECK_ARG(group_mode >= 0 && group_mode <= 2, "fxc_plan_paths: group mode 0, 1 or 2");
  FXC_CHECK_ARG(!use_azimuth || (az_ranges && nranges >= 1),
                "fxc_plan_paths: azimuth filter needs ranges");
  *out_set = nullptr;
  Consts C;
  C.pi = M_PI;
  C.d2r = C.pi / 180.0;
  C.r2d = 180.0 / C.pi;
  C.ec2 = 2.0 * kF - kF * kF;
  C.one_minus_ec2 = 1.0 - C.ec2;
  C.e_prime = C.ec2 / C.one_minus_ec2;
  PlanConfig cfg{allow_autocorr, autocorr_only, group_mode, use_distance, min_km, max_km,
                 use_azimuth, {}};
  if (use_azimuth) cfg.ranges.assign(az_ranges, az_ranges + 2 * nranges);
  std::vector<Node> nodes((size_t)nnodes);
  for (int i = 0; i < nnodes; ++i) {
    nodes[i].geo = has_geo[i] != 0;
    nodes[i].lat = lat[i];
    nodes[i].lon = lon[i];
    if (nodes[i].geo) node_terms(nodes[i], C.d2r, C.one_minus_ec2);
  }
  std::vector<std::vector<Path>> rows((size_t)nnodes);
  std::atomic<int> next{0};
  std::atomic<long long> checks{0};
  auto work = [&]() {
    long long my_checks = 0;
    for (int i; (i = next.fetch_add(1)) < nnodes;) {
      std::vector<Path> &out = rows[(size_t)i];
      const int j0 = cfg.autocorr_only ? i : (cfg.allow_autocorr ? i : i + 1);
      const int j1 = cfg.autocorr_only ? i + 1 : nnodes;
      for (int j = j0; j < j1; ++j) {
        ++my_checks;
        const bool autoc = i == j;  // nodes are unique, so equal keys mean equal index
        if (autoc && !cfg.allow_autocorr) continue;
        if (cfg.autocorr_only && !autoc) continue;
        if (!autoc && cfg.group_mode == 1 && group[i] != group[j]) continue;
        if (!autoc && cfg.group_mode == 2 && group[i] == group[j]) continue;
        Path p{i, j, 0, {0, 0, 0, 0}};
        if (!nodes[i].geo || !nodes[j].geo) {
          if (cfg.use_distance || cfg.use_azimuth) continue;
        } else {
          geometry(nodes[i], nodes[j], C, p.g);
          p.has_geo = 1;
        }
        if (cfg.use_distance) {
          if (!p.has_geo) continue;
          if (p.g[3] < cfg.min_km || p.g[3] > cfg.max_km) continue;
        }
        if (cfg.use_azimuth) {
          if (!p.has_geo) continue;
          const int nr = (int)cfg.ranges.size() / 2;
          if (!(in_ranges(p.g[1], cfg.ranges.data(), nr) ||
                in_ranges(p.g[2], cfg.ranges.data(), nr)))
            continue;
        }
        out.push_back(p);
      }
    }
    checks += my_checks;
  };
  int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, 64));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
  PathSet *set = new PathSet();
  size_t total = 0;
  for (const auto &r : rows) total += r.size();
  set->paths.reserve(total);
  for (auto &r : rows) set->paths.insert(set->paths.end(), r.begin(), r.end());
  set->checks = checks.load();
  *out_set = set;
  *npaths = (long long)total;
  *nchecks = set->checks;
  return FXC_OK;
}

extern "C" int fxc_plan_paths_fetch(void *set_ptr, int *src, int *rec, unsigned char *has_geo,
                                    double *geo) {
  FXC_CHECK_ARG(set_ptr && src && rec && has_geo && geo, "fxc_plan_paths_fetch: invalid arguments");
  const PathSet *set = (const PathSet *)set_ptr;
  for (size_t k = 0; k < set->paths.size(); ++k) {
    const Path &p = set->paths[k];
    src[k] = p.src;
    rec[k] = p.rec;
    has_geo[k] = p.has_geo;
    for (int q = 0; q < 4; ++q) geo[4 * k + q] = p.g[q];
  }
  return FXC_OK;
}

extern "C" int fxc_plan_paths_free(void *set_ptr) {
  delete (PathSet *)set_ptr;
  return FXC_OK;
}

namespace {

// csv.writer's field with a tab delimiter and minimal quoting
void put_field(std::string &out, const char *v) {
  if (!std::strpbrk(v, "\t\"\r\n")) {
    out += v;
    return;
  }
  out += '"';
  for (const char *c = v; *c; ++c) {
    if (*c == '"') out += '"';
    out += *c;
  }
  out += '"';
}

// Python's f"{v:.7g}" (glibc's %.7g rounds the exact binary value the same way)
void put_g7(std::string &out, double v) {
  char buf[40];
  std::snprintf(buf, sizeof(buf), "%.7g", v);
  out += buf;
}

}  // namespace

extern "C" int fxc_plan_paths_write(void *set_ptr, int nnodes, const int *node_id,
                                    const char *const *node_text, const char *tsv_path,
                                    const char *ids_path, int nthreads) {
  FXC_CHECK_ARG(set_ptr && node_id && node_text && tsv_path && ids_path && nnodes >= 0,
                "fxc_plan_paths_write: invalid arguments");
  const PathSet *set = (const PathSet *)set_ptr;
  const size_t n = set->paths.size();
  for (const Path &p : set->paths)
    FXC_CHECK_ARG(p.src < nnodes && p.rec < nnodes, "fxc_plan_paths_write: node table too short");
  int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
  nt = std::max(1, std::min(nt, 64));
  const size_t per = (n + (size_t)nt - 1) / (size_t)nt;
  std::vector<std::string> tsv((size_t)nt), ids((size_t)nt);
  auto fmt = [&](int t) {
    const size_t a = std::min(n, (size_t)t * per), b = std::min(n, a + per);
    std::string &o = tsv[(size_t)t], &q = ids[(size_t)t];
    o.reserve((b - a) * 96);
    q.reserve((b - a) * 9);
    char idb[32];
    for (size_t k = a; k < b; ++k) {
      const Path &p = set->paths[k];
      const int s = node_id[p.src], r = node_id[p.rec];
      std::snprintf(idb, sizeof(idb), "%08lld", (long long)s * 10000LL + r);
      o += idb;
      q += idb;
      q += '\n';
      std::snprintf(idb, sizeof(idb), "\t%04d\t%04d\t", s, r);
      o += idb;
      put_field(o, node_text[p.src]);
      o += '\t';
      put_field(o, node_text[p.rec]);
      // great circle, distance, azimuth, back azimuth
      const int order[4] = {0, 3, 1, 2};
      for (int c = 0; c < 4; ++c) {
        o += '\t';
        if (p.has_geo) put_g7(o, p.g[order[c]]);
      }
      o += "\r\n";
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(fmt, t);
  fmt(0);
  for (auto &th : pool) th.join();
  FILE *f = std::fopen(tsv_path, "wb");
  FXC_CHECK_ARG(f, "fxc_plan_paths_write: cannot open %s", tsv_path);
  static const char kHead[] =
      "path_id\tsrc_nsl_id\trec_nsl_id\tsrc_nsl\trec_nsl\tgreat_circle_deg\tdistance_km\t"
      "azimuth_deg\tback_azimuth_deg\r\n";
  bool ok = std::fwrite(kHead, 1, sizeof(kHead) - 1, f) == sizeof(kHead) - 1;
  for (const auto &chunk : tsv) ok = ok && std::fwrite(chunk.data(), 1, chunk.size(), f) == chunk.size();
  ok = (std::fclose(f) == 0) && ok;
  FXC_CHECK_ARG(ok, "fxc_plan_paths_write: write to %s failed", tsv_path);
  f = std::fopen(ids_path, "wb");
  FXC_CHECK_ARG(f, "fxc_plan_paths_write: cannot open %s", ids_path);
  ok = true;
  for (const auto &chunk : ids) ok = ok && std::fwrite(chunk.data(), 1, chunk.size(), f) == chunk.size();
  ok = (std::fclose(f) == 0) && ok;
  FXC_CHECK_ARG(ok, "fxc_plan_paths_write: write to %s failed", ids_path);
  return FXC_OK;
}

