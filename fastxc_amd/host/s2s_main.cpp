// sac2spec: SAC windows -> whitened spectra in StepPack v3, on one MI355X.
//
// Command line, output layout and markers follow the reference executable
// (native/sac2spec/arguproc.c:200-420, native/sac2spec/cuda.main.cu:73-536,
// output_layout.cu, timestamp_tracker.cu, progress.cu):
//   sac2spec -I sac_index.tsv -O ws -C nch -B filter.txt -L sac/win/shift/lag
//            [-G ids] [-W whiten/phase] [-N norm | --normalize N]
//            [-Q steps,...,-1] [-T threads] [-M MiB,...] [--debug] [--dry-run]
// writes ws/stepack/w000.b%06zu.{stepack,tsv} (one file per batch of station
// groups), ws/stepack/<timestamp>/_SUCCESS when every group of a timestamp is
// written, ws/stepack/_SUCCESS at the end, and ws/progress/sac2spec_progress.tsv.
//
// Per batch: the SAC files are read by -T threads into pinned memory
// (min(npts_file, npts) samples, zero padded; the sampling interval must
// match the first file's within 1e-5 relative / 1e-6 absolute), uploaded once,
// processed by fxc_s2s_process (include/fastxc_amd.h) and written back as
// one StepPack record.
#include "fxc_host.hpp"
#include "s2s_plan.hpp"
#include "stepack.hpp"

#include "fastxc_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

using namespace fxh;

namespace {

struct Args {
  std::string input, out, filter, gpu_ids = "0", mem_caps;
  int num_ch = 3, threads = 1, whiten = 0, phase_only = 0, normalize = 0;
  float sac_len = 0.f, win_len = 0.f, shift_len = 0.f, lag = 0.f;
  bool window_set = false, debug = false, dry_run = false;
  std::vector<int> skip;
};

void usage() {
  fprintf(stderr,
          "Usage: sac2spec -I sac_index.tsv -O output_root -C num_ch -B filter.txt\n"
          "                -L sac_len/win_len/shift_len/xcorr_lag [options]\n"
          "  -I  TSV: timestamp nsl_id network station location component sac_path\n"
          "  -O  output root (writes stepack/ and progress/)\n"
          "  -C  channels per station group (default 3)\n"
          "  -B  Butterworth filter file (band 0 = broad band)\n"
          "  -L  seconds: SAC length / window / shift / max correlation lag\n"
          "  -G  GPU list; the first id is used (one GPU per process)\n"
          "  -W  whiten_type/output_phase_only: 0 none, 1 before, 2 after, 3 both; 0|1\n"
          "  -N  normalisation: 0 none, 1 RUN-ABS-MF, 2 one-bit, 3 RUN-ABS\n"
          "  -Q  window indices to skip, comma separated, ended by -1\n"
          "  -T  SAC reader threads (default 1)\n"
          "  -M  device memory cap in MiB for the first GPU\n"
          "  --normalize N   same as -N\n"
          "  --debug         synchronous processing\n"
          "  --dry-run       print the derived plan (lengths, taper bins, filter\n"
          "                  responses) and exit without touching the GPU\n");
}

bool parse_float(const char *s, float *out) {
  char *end = nullptr;
  errno = 0;
  const float v = strtof(s, &end);
  if (errno || end == s || *end) return false;
  *out = v;
  return true;
}

bool parse_int(const char *s, int *out) {
  char *end = nullptr;
  errno = 0;
  const long v = strtol(s, &end, 10);
  if (errno || end == s || *end || v < -2147483647L || v > 2147483647L) return false;
  *out = (int)v;
  return true;
}

bool fail(const char *msg) {
  fprintf(stderr, "Error: %s\n", msg);
  return false;
}

bool parse_args(int argc, char **argv, Args *a, int *code) {
  *code = 1;
  if (argc < 2) {
    usage();
    return false;
  }
  for (int i = 1; i < argc; ++i) {
    const std::string o = argv[i];
    if (o == "-h" || o == "--help") {
      usage();
      *code = 0;
      return false;
    }
    auto need = [&]() -> const char * {
      if (i + 1 >= argc) {
        fprintf(stderr, "Error: Missing value for %s\n", o.c_str());
        usage();
        return nullptr;
      }
      return argv[++i];
    };
    if (o == "--normalize" || o == "--normalize-type") {
      const char *v = need();
      if (!v) return false;
      a->normalize = atoi(v);
      continue;
    }
    if (o == "--debug") {
      a->debug = true;
      continue;
    }
    if (o == "--dry-run") {
      a->dry_run = true;
      continue;
    }
    if (o.compare(0, 2, "--") == 0) {
      fprintf(stderr, "Error: Unknown option %s\n", o.c_str());
      usage();
      return false;
    }
    if (o.size() != 2 || o[0] != '-') {
      fprintf(stderr, "Error: Invalid option %s\n", o.c_str());
      usage();
      return false;
    }
    const char *v = need();
    if (!v) return false;
    switch (o[1]) {
      case 'I': a->input = v; break;
      case 'O': a->out = v; break;
      case 'C': a->num_ch = atoi(v); break;
      case 'B': a->filter = v; break;
      case 'G': a->gpu_ids = v; break;
      case 'M': a->mem_caps = v; break;
      case 'T': a->threads = atoi(v); break;
      case 'N': a->normalize = atoi(v); break;
      case 'L': {
        std::vector<std::string> p;
        std::string s = v, cur;
        for (char c : s) {
          if (c == '/') {
            p.push_back(cur);
            cur.clear();
          } else {
            cur += c;
          }
        }
        p.push_back(cur);
        if (p.size() != 4 || !parse_float(p[0].c_str(), &a->sac_len) ||
            !parse_float(p[1].c_str(), &a->win_len) || !parse_float(p[2].c_str(), &a->shift_len) ||
            !parse_float(p[3].c_str(), &a->lag))
          return fail("Invalid -L value; expected sac_len/win_len/shift_len/xcorr_lag");
        a->window_set = true;
        break;
      }
      case 'W': {
        const std::string s = v;
        const size_t sl = s.find('/');
        if (sl == std::string::npos || s.find('/', sl + 1) != std::string::npos ||
            !parse_int(s.substr(0, sl).c_str(), &a->whiten) ||
            !parse_int(s.substr(sl + 1).c_str(), &a->phase_only))
          return fail("Invalid -W value; expected whiten_type/output_phase_only");
        break;
      }
      case 'Q': {
        std::string s = v, tok;
        s += ",";
        for (char c : s) {
          if (c != ',') {
            tok += c;
            continue;
          }
          int val = 0;
          if (!parse_int(tok.c_str(), &val)) return fail("Invalid -Q entry");
          tok.clear();
          if (val == -1) break;
          if (val < 0) return fail("Invalid -Q entry; steps must be >= 0");
          if (a->skip.size() >= 100) return fail("Too many skip steps; max is 100");
          a->skip.push_back(val);
        }
        break;
      }
      default:
        fprintf(stderr, "Error: Unknown option -%c\n", o[1]);
        usage();
        return false;
    }
  }
  if (a->input.empty() || a->out.empty() || a->filter.empty() || !a->window_set ||
      a->win_len <= 0.f) {
    if (!a->window_set) fprintf(stderr, "Error: -L sac_len/win_len/shift_len/xcorr_lag is required\n");
    usage();
    return false;
  }
  if (a->shift_len <= 0.f) return fail("-L shift_len must be > 0");
  if (a->sac_len <= 0.f) return fail("-L sac_len must be > 0");
  if (a->num_ch < 1) return fail("-C must be >= 1");
  if (a->threads < 1) return fail("-T must be >= 1");
  if (a->lag < 0.f) return fail("-L xcorr_lag must be >= 0");
  if (a->whiten < 0 || a->whiten > 3) return fail("-W whiten_type must be 0, 1, 2, or 3");
  if (a->phase_only != 0 && a->phase_only != 1) return fail("-W output_phase_only must be 0 or 1");
  if (a->normalize < 0 || a->normalize > 3) return fail("-N normalize type must be 0, 1, 2, or 3");
  return true;
}

// ----------------------------------------------------------- sac_index.tsv --
struct IndexRow {
  StepackNode node;
  std::string path;
};

// trims, and turns inner blank runs into '_' (io/sac_index.c copyCleanField)
std::string clean_field(const std::string &s) {
  std::string o;
  bool space = false;
  for (char c : trim(s)) {
    if (isspace((unsigned char)c)) {
      if (!o.empty() && !space) o += '_';
      space = true;
    } else {
      o += c;
      space = false;
    }
  }
  while (!o.empty() && o.back() == '_') o.pop_back();
  return o;
}

// Rows grouped by consecutive (timestamp, nsl_id); a group is kept when it
// has num_ch rows (or num_ch is 1), otherwise skipped with a warning.
bool read_sac_index(const std::string &path, int num_ch, std::vector<IndexRow> *rows) {
  FILE *fp = fopen(path.c_str(), "r");
  if (!fp) {
    FXH_LOG_ERROR("sac_index_open_failed", "path=\"%s\"", path.c_str());
    return false;
  }
  std::vector<IndexRow> group;
  auto flush = [&]() {
    if (group.empty()) return;
    if (num_ch == 1 || (int)group.size() == num_ch) {
      rows->insert(rows->end(), group.begin(), group.end());
    } else {
      FXH_LOG_WARN("sac_index_group_skipped",
                   "timestamp=%s nsl_id=%s path=\"%s\" expected_components=%d got_components=%zu",
                   group[0].node.timestamp.c_str(), group[0].node.nsl_id.c_str(), path.c_str(),
                   num_ch, group.size());
    }
    group.clear();
  };
  char line[8192];
  int line_no = 0;
  bool ok = true;
  while (fgets(line, sizeof(line), fp)) {
    ++line_no;
    const std::string t = trim(line);
    if (t.empty() || t[0] == '#' || t.compare(0, 9, "timestamp") == 0 || t.compare(0, 6, "nsl_id") == 0)
      continue;
    const auto f = split_tab(t);
    if (f.size() < 7) {
      FXH_LOG_ERROR("sac_index_row_malformed", "path=\"%s\" line=%d", path.c_str(), line_no);
      ok = false;
      break;
    }
    IndexRow r;
    r.node.timestamp = clean_field(f[0]);
    r.node.nsl_id = clean_field(f[1]);
    r.node.network = clean_field(f[2]);
    r.node.station = clean_field(f[3]);
    r.node.location = clean_field(f[4]);
    r.node.component = clean_field(f[5]);
    std::string p = trim(f[6]);
    for (size_t k = 7; k < f.size(); ++k) p += "\t" + f[k];
    r.path = trim(p);
    if (r.node.timestamp.empty() || r.node.nsl_id.empty() || r.node.network.empty() ||
        r.node.station.empty() || r.node.location.empty() || r.node.component.empty() ||
        r.path.empty()) {
      FXH_LOG_ERROR("sac_index_row_missing_field", "path=\"%s\" line=%d", path.c_str(), line_no);
      ok = false;
      break;
    }
    char *end = nullptr;
    const long id = strtol(r.node.nsl_id.c_str(), &end, 10);
    if (end == r.node.nsl_id.c_str() || *end || id < 0 || id > 9999) {
      FXH_LOG_ERROR("sac_index_nsl_id_invalid", "nsl_id=\"%s\" path=\"%s\"", r.node.nsl_id.c_str(),
                    r.path.c_str());
      ok = false;
      break;
    }
    r.node.gnsl_id = (uint32_t)id;
    if (!group.empty() && (group[0].node.nsl_id != r.node.nsl_id ||
                           group[0].node.timestamp != r.node.timestamp))
      flush();
    group.push_back(r);
  }
  fclose(fp);
  if (!ok) return false;
  flush();
  if (rows->empty()) {
    FXH_LOG_ERROR("sac_index_empty", "path=\"%s\"", path.c_str());
    return false;
  }
  return true;
}

// timestamp text -> directory leaf: [0-9A-Za-z._:-] kept, others '-', trailing '-' cut
std::string safe_timestamp_leaf(const std::string &ts) {
  std::string o;
  for (char c : ts)
    o += (isalnum((unsigned char)c) || c == '.' || c == '_' || c == ':' || c == '-') ? c : '-';
  while (!o.empty() && o.back() == '-') o.pop_back();
  return o.empty() ? "timestamp" : o;
}

#define HIPX(call)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      FXH_LOG_ERROR("hip_error", "call=\"%s\" error=\"%s\"", #call, hipGetErrorString(e_)); \
      return 1;                                                                              \
    }                                                                                        \
  } while (0)

int run(const Args &a) {
  const std::string stepack_root = path_join(a.out, "stepack");
  const std::string progress_dir = path_join(a.out, "progress");
  if (mkdir_p(stepack_root) != 0 || mkdir_p(progress_dir) != 0) {
    FXH_LOG_ERROR("output_layout_failed", "root=\"%s\"", a.out.c_str());
    return 1;
  }
  if (!a.dry_run && dir_has_entries(stepack_root)) {
    FXH_LOG_ERROR("stepack_root_not_empty", "path=\"%s\"", stepack_root.c_str());
    return 1;
  }
  std::vector<IndexRow> rows;
  if (!read_sac_index(a.input, a.num_ch, &rows)) return 1;
  if (rows.size() % (size_t)a.num_ch != 0) {
    FXH_LOG_ERROR("sac_index_channel_count_invalid", "rows=%zu num_ch=%d", rows.size(), a.num_ch);
    return 1;
  }
  const size_t ngroups = rows.size() / (size_t)a.num_ch;
  SacHeader h0;
  if (read_sac_header(rows[0].path, &h0) != 0) return 1;
  const float delta = h0.delta;
  if (!(delta > 0.f)) {
    FXH_LOG_ERROR("sac_delta_invalid", "path=\"%s\" delta=%.9g", rows[0].path.c_str(), (double)delta);
    return 1;
  }
  const std::vector<Butterworth> filters = read_filters(a.filter);
  S2sLengths L;
  if (!derive_lengths(a.sac_len, delta, a.win_len, a.shift_len, a.lag, filters, &L)) return 1;
  std::vector<int> valid;
  for (int s = 0; s < L.nstep; ++s)
    if (std::find(a.skip.begin(), a.skip.end(), s) == a.skip.end()) valid.push_back(s);
  if (valid.empty()) {
    FXH_LOG_ERROR("all_steps_skipped", "nstep=%d skip_count=%zu", L.nstep, a.skip.size());
    return 1;
  }
  const int nsteps = (int)valid.size();
  FXH_LOG_INFO("plan", "npts=%d delta=%.9g seg=%d shift=%d nstep=%d valid=%d nfft_filter=%d nfft_out=%d nspec=%d bands=%zu",
               L.npts, (double)delta, L.seg, L.shift, L.nstep, nsteps, L.nfft_filter, L.nfft_out,
               L.nspec, filters.size());
  const int half_f = L.nfft_filter / 2 + 1;
  std::vector<float> resp((size_t)filters.size() * half_f), lows(filters.size());
  for (size_t i = 0; i < filters.size(); ++i) {
    const auto r = power_response(filters[i], L.nfft_filter);
    std::copy(r.begin(), r.end(), resp.begin() + i * half_f);
    lows[i] = filters[i].freq_low;
  }

  if (a.dry_run) {
    printf("groups\t%zu\nnpts\t%d\ndelta\t%.9g\nseg\t%d\nshift\t%d\nnstep\t%d\nvalid_steps\t",
           ngroups, L.npts, (double)delta, L.seg, L.shift, L.nstep);
    for (size_t i = 0; i < valid.size(); ++i) printf(i ? ",%d" : "%d", valid[i]);
    printf("\nnfft_filter\t%d\nnfft_out\t%d\nnspec\t%d\ndf_out\t%.9g\nidx\t%d,%d,%d,%d\nbands\t%zu\n",
           L.nfft_filter, L.nfft_out, L.nspec, (double)L.df_out, L.idx[0], L.idx[1], L.idx[2],
           L.idx[3], filters.size());
    for (size_t i = 0; i < filters.size(); ++i) {
      printf("band%zu\t%.9g/%.9g\t", i, (double)filters[i].freq_low, (double)filters[i].freq_high);
      for (int k = 0; k < half_f; ++k) printf(k ? ",%.9g" : "%.9g", (double)resp[i * half_f + k]);
      printf("\n");
    }
    return 0;
  }
  const int dev = atoi(a.gpu_ids.c_str());
  HIPX(hipSetDevice(dev));
  hipStream_t st;
  HIPX(hipStreamCreate(&st));
  fxc_s2s_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.npts = L.npts;
  cfg.seg = L.seg;
  cfg.shift = L.shift;
  cfg.nfft_filter = L.nfft_filter;
  cfg.nfft_out = L.nfft_out;
  cfg.num_ch = a.num_ch;
  cfg.delta = delta;
  cfg.freq_low = filters[0].freq_low;
  for (int i = 0; i < 4; ++i) cfg.idx[i] = L.idx[i];
  cfg.whiten = a.whiten;
  cfg.normalize = a.normalize;
  cfg.phase_only = a.phase_only;
  cfg.filter_count = (int)filters.size();
  cfg.responses = resp.data();
  cfg.freq_lows = lows.data();
  fxc_s2s_plan *plan = nullptr;
  if (fxc_s2s_plan_create(&plan, &cfg) != FXC_OK) {
    FXH_LOG_ERROR("s2s_plan_failed", "error=\"%s\"", fxc_last_error());
    return 1;
  }

  // groups per batch: bounded by device memory (traces + spectra, x2 for the
  // kernels' workspaces) and by 256 groups of pinned host memory
  size_t mem_free = 0, mem_total = 0;
  HIPX(hipMemGetInfo(&mem_free, &mem_total));
  size_t budget = (size_t)(0.8 * (double)mem_free);
  const long long cap_mib = a.mem_caps.empty() ? 0 : atoll(a.mem_caps.c_str());
  if (cap_mib > 0) budget = std::min(budget, (size_t)cap_mib << 20);
  const size_t per_group = (size_t)a.num_ch * ((size_t)L.npts * 4 + (size_t)nsteps * L.nspec * 8) * 2;
  size_t batch_groups = std::max<size_t>(1, std::min<size_t>(budget / per_group, 256));
  batch_groups = std::min(batch_groups, ngroups);

  const size_t max_files = batch_groups * (size_t)a.num_ch;
  float *h_traces = nullptr, *d_traces = nullptr;
  std::complex<float> *h_spec = nullptr;
  float *d_spec = nullptr;
  HIPX(hipHostMalloc((void **)&h_traces, max_files * (size_t)L.npts * 4));
  HIPX(hipHostMalloc((void **)&h_spec, max_files * (size_t)nsteps * L.nspec * 8));
  HIPX(hipMalloc((void **)&d_traces, max_files * (size_t)L.npts * 4));
  HIPX(hipMalloc((void **)&d_spec, max_files * (size_t)nsteps * L.nspec * 8));

  Progress prog;
  const std::string prog_path = path_join(progress_dir, "sac2spec_progress.tsv");
  prog.init(prog_path.c_str());
  prog.set_rows({{"overall", "RUNNING", 0, rows.size(), "sac_rows",
                  "gpu_id=-1 last_batch_rows=0 start_group=0 group_count=0"}});
  std::map<std::string, size_t> ts_total, ts_done;
  for (size_t g = 0; g < ngroups; ++g) ++ts_total[rows[g * a.num_ch].node.timestamp];

  int rc = 0;
  size_t batch_seq = 0, done_rows = 0;
  const float df = L.df_out;
  for (size_t g0 = 0; g0 < ngroups && rc == 0; g0 += batch_groups, ++batch_seq) {
    const size_t gcount = std::min(batch_groups, ngroups - g0);
    const size_t nfiles = gcount * (size_t)a.num_ch;
    const size_t f0 = g0 * (size_t)a.num_ch;
    std::vector<StepackNode> nodes(nfiles);
    std::atomic<size_t> next(0);
    std::atomic<bool> bad(false);
    auto reader = [&]() {
      for (;;) {
        const size_t i = next.fetch_add(1);
        if (i >= nfiles || bad) return;
        const IndexRow &r = rows[f0 + i];
        SacHeader h;
        std::vector<float> buf;
        if (read_sac(r.path, &h, &buf, (size_t)L.npts) != 0) {
          bad = true;
          return;
        }
        const float tol = std::max(std::fabs(delta) * 1.0e-5f, 1.0e-6f);
        if (std::fabs(h.delta - delta) > tol) {
          FXH_LOG_ERROR("sac_delta_mismatch", "path=\"%s\" expected_delta=%.9g source_delta=%.9g",
                        r.path.c_str(), (double)delta, (double)h.delta);
          bad = true;
          return;
        }
        if (h.npts != L.npts)
          FXH_LOG_WARN("sac_npts_adjusted", "path=\"%s\" source_npts=%d target_npts=%d action=%s",
                       r.path.c_str(), h.npts, L.npts,
                       h.npts > L.npts ? "truncated" : "zero_padded");
        memcpy(h_traces + i * (size_t)L.npts, buf.data(), (size_t)L.npts * 4);
        nodes[i] = r.node;
        nodes[i].stla = h.stla;
        nodes[i].stlo = h.stlo;
      }
    };
    std::vector<std::thread> pool;
    const int nth = std::max(1, std::min<int>(a.threads, (int)nfiles));
    for (int t = 0; t < nth; ++t) pool.emplace_back(reader);
    for (auto &t : pool) t.join();
    if (bad) {
      FXH_LOG_ERROR("worker_sac_load_failed", "gpu=%d start_group=%zu count=%zu", dev, g0, gcount);
      rc = 1;
      break;
    }
    HIPX(hipMemcpyAsync(d_traces, h_traces, nfiles * (size_t)L.npts * 4, hipMemcpyHostToDevice, st));
    if (fxc_s2s_process(plan, d_traces, (int)nfiles, valid.data(), nsteps, d_spec, st) != FXC_OK) {
      FXH_LOG_ERROR("s2s_process_failed", "error=\"%s\"", fxc_last_error());
      rc = 1;
      break;
    }
    HIPX(hipMemcpyAsync(h_spec, d_spec, nfiles * (size_t)nsteps * L.nspec * 8,
                        hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    if (stepack_write_batch(stepack_root, 0, batch_seq, g0, gcount, nodes, nsteps, L.nspec, delta,
                            df, h_spec) != 0) {
      rc = 1;
      break;
    }
    done_rows += nfiles;
    char detail[160];
    snprintf(detail, sizeof(detail), "gpu_id=%d last_batch_rows=%zu start_group=%zu group_count=%zu",
             dev, nfiles, g0, gcount);
    prog.update("overall", "RUNNING", done_rows, rows.size(), "sac_rows", detail);
    for (size_t g = g0; g < g0 + gcount; ++g) {
      const std::string &ts = rows[g * a.num_ch].node.timestamp;
      if (++ts_done[ts] == ts_total[ts]) {
        const std::string dir = path_join(stepack_root, safe_timestamp_leaf(ts));
        FILE *fp = (mkdir_p(dir) == 0) ? fopen(path_join(dir, "_SUCCESS").c_str(), "wb") : nullptr;
        if (!fp) {
          FXH_LOG_ERROR("timestamp_success_write_failed", "timestamp=%s", ts.c_str());
          rc = 1;
          break;
        }
        fprintf(fp, "timestamp\t%s\ngroups\t%zu\n", ts.c_str(), ts_total[ts]);
        fclose(fp);
      }
    }
    FXH_LOG_INFO("batch_done", "batch_seq=%zu start_group=%zu group_count=%zu rows=%zu", batch_seq,
                 g0, gcount, nfiles);
  }
  fxc_s2s_plan_destroy(plan);
  (void)hipHostFree(h_traces);
  (void)hipHostFree(h_spec);
  (void)hipFree(d_traces);
  (void)hipFree(d_spec);
  (void)hipStreamDestroy(st);
  if (rc == 0 && write_marker(path_join(stepack_root, "_SUCCESS")) != 0) {
    FXH_LOG_ERROR("stepack_success_write_failed", "path=\"%s\"", stepack_root.c_str());
    rc = 1;
  }
  prog.update("overall", rc == 0 ? "DONE" : "FAILED", rc == 0 ? rows.size() : done_rows,
              rows.size(), "sac_rows", "gpu_id=-1 last_batch_rows=0 start_group=0 group_count=0");
  return rc;
}

}  // namespace

int main(int argc, char **argv) {
  Args a;
  int code = 1;
  if (!parse_args(argc, argv, &a, &code)) return code;
  FXH_LOG_INFO("run_start", "input=\"%s\" output=\"%s\" num_ch=%d threads=%d", a.input.c_str(),
               a.out.c_str(), a.num_ch, a.threads);
  return run(a);
}
