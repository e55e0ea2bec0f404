// xc_fast: windowed cross-correlation of every allowed station pair, one
// timestamp at a time, on one MI355X.
//
// Command line and outputs follow the reference executable
// (native/xc/arguproc.c:165-261, native/xc/cuda.main.cu:345-421):
//   xc_fast -I <stepack_dir|stepack.tsv> -P allowed_paths.tsv -O out -C sec
//           [-G ids] [-M MiB,...] [-J depth] [--progress file] [--debug]
// writes out/xcpack/<ts>.iAAAAAA.jBBBBBB.wWWW.pPPP.{xcpack,tsv} (records
// [SAC header][2H+1 float], parts cut at 4 GiB), out/xcpack/<ts>.done after
// each timestamp and out/xcpack/_SUCCESS at the end, and the progress sidecar.
// Extensions: --precision fp32|bf16x3|fp16|fp16x3|plain (default fp32, the
// reference's arithmetic, or the FASTXC_XC_PRECISION environment value),
// --block-files n (block size; the reference derives it from device memory,
// at most 4096), --dry-run (print jobs and tasks, no GPU) and
// --sourcepack-dir d (each finished timestamp's SourcePack index built on a
// side thread, as the reference's Python materializer does by polling the
// .done markers; fastxc/operators/sourcepack/builder.py:26-160).
//
// Per timestamp: the StepPack payload [nstep][rows][nspec] is read into
// pinned memory and uploaded once; each job's tasks go to fxc_xc_correlate
// (include/fastxc_amd.h); the correlations come back to pinned buffers and a
// writer thread turns them into xcpack records while the GPU runs the next
// job or timestamp.
#include "fxc_host.hpp"
#include "sourcepack.hpp"
#include "stepack.hpp"
#include "xc_tasks.hpp"

#include "fastxc_amd.h"

#include <hip/hip_runtime.h>

#include <inttypes.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace fxh;

namespace {

constexpr size_t kMaxBlockFiles = 4096;       // block size ceiling of the reference's planner
constexpr uint64_t kMaxPackBytes = 4ull << 30;  // xcpack part size

struct Args {
  std::string input, paths, out, progress, gpu_ids = "0", mem_caps;
  std::string sourcepack_dir, sourcepack_progress;
  bool sourcepack_sort = true;
  float cclength = 0.f;
  int lazy_depth = 1;
  size_t block_files = 0;  // 0: min(rows, 4096)
  bool debug = false, dry_run = false;
  std::string precision = "fp32";
};

void usage() {
  fprintf(stderr,
          "\nUsage:\n"
          "  xc_fast -I <stepack_dir|stepack.tsv> -P allowed_paths.tsv -O out -C sec [options]\n\n"
          "Input:\n"
          "  -I <path>            Stepack directory or stepack TSV sidecar\n"
          "  -P <file>            Canonical allowed path table\n"
          "  -O <dir>             Output root\n"
          "  -C <sec>             Half CC length\n\n"
          "Runtime:\n"
          "  -G <ids>             GPU list; the first id is used (one GPU per process)\n"
          "  -M <MiB,...>         Device memory cap for the first GPU; 0 means auto\n"
          "  -J <num>             Write queue depth (default 1, 0 writes synchronously)\n"
          "  --precision <p>      fp32 | bf16x3 | fp16 | fp16x3 | plain (default: FASTXC_XC_PRECISION\n"
          "                       if set, else fp32)\n"
          "  --block-files <n>    Station rows per block (default min(rows, 4096))\n"
          "  --dry-run            Print the jobs and tasks of every timestamp and exit\n"
          "  --debug              Write synchronously\n"
          "  --progress <file>    Write progress sidecar TSV\n\n"
          "Source index:\n"
          "  --sourcepack-dir <d> Build <d>/<timestamp>/sourcepack_index.tsv as each\n"
          "                       timestamp is done, and <d>/_SUCCESS at the end\n"
          "  --sourcepack-progress <file>  Progress sidecar of the index build\n"
          "  --sourcepack-keep-order       Keep the xcpack order inside each index\n\n"
          "Logging:\n"
          "  FASTXC_LOG_LEVEL=ERROR|WARN|INFO|DEBUG (default INFO)\n\n"
          "Output:\n"
          "  Writes <output>/xcpack/*.xcpack + *.tsv sidecars.\n\n");
}

bool is_option(const char *s) { return s && s[0] == '-' && s[1] != '\0'; }

bool parse_args(int argc, char **argv, Args *a, int *exit_code) {
  *exit_code = 1;
  // deployment default for the contraction form (the stage command line
  // built by the Python driver has no precision flag)
  if (const char *env = getenv("FASTXC_XC_PRECISION"))
    if (*env) a->precision = env;
  if (argc <= 1) {
    usage();
    return false;
  }
  for (int i = 1; i < argc; ++i) {
    const std::string o = argv[i];
    auto value = [&](const char *opt) -> const char * {
      if (i + 1 >= argc || is_option(argv[i + 1])) {
        FXH_LOG_ERROR("argument_value_missing", "option=\"%s\"", opt);
        return nullptr;
      }
      return argv[++i];
    };
    const char *v = nullptr;
    if (o == "-h" || o == "--help") {
      usage();
      *exit_code = 0;
      return false;
    } else if (o == "-I") {
      if (!(v = value("-I"))) break;
      a->input = v;
    } else if (o == "-O") {
      if (!(v = value("-O"))) break;
      a->out = v;
    } else if (o == "-C") {
      if (!(v = value("-C"))) break;
      a->cclength = (float)atof(v);
    } else if (o == "-G") {
      if (!(v = value("-G"))) break;
      a->gpu_ids = v;
    } else if (o == "-M") {
      if (!(v = value("-M"))) break;
      a->mem_caps = v;
    } else if (o == "-P") {
      if (!(v = value("-P"))) break;
      a->paths = v;
    } else if (o == "-J") {
      if (!(v = value("-J"))) break;
      a->lazy_depth = atoi(v);
    } else if (o == "--progress" || o == "--progress-file") {
      if (!(v = value(o.c_str()))) break;
      a->progress = v;
    } else if (o == "--precision") {
      if (!(v = value("--precision"))) break;
      a->precision = v;
    } else if (o == "--block-files") {
      if (!(v = value("--block-files"))) break;
      a->block_files = (size_t)atoll(v);
    } else if (o == "--sourcepack-dir") {
      if (!(v = value("--sourcepack-dir"))) break;
      a->sourcepack_dir = v;
    } else if (o == "--sourcepack-progress") {
      if (!(v = value("--sourcepack-progress"))) break;
      a->sourcepack_progress = v;
    } else if (o == "--sourcepack-keep-order") {
      a->sourcepack_sort = false;
    } else if (o == "--debug") {
      a->debug = true;
    } else if (o == "--dry-run") {
      a->dry_run = true;
    } else {
      FXH_LOG_ERROR("argument_unknown", "option=\"%s\"", o.c_str());
      usage();
      return false;
    }
    continue;
  }
  if (a->out.empty() || a->cclength <= 0.f || a->paths.empty()) {
    FXH_LOG_ERROR("required_argument_missing", "required=\"-O,-C,-P\"");
    usage();
    return false;
  }
  if (a->input.empty()) {
    FXH_LOG_ERROR("timestamp_input_missing", "required=\"-I\"");
    usage();
    return false;
  }
  if (a->precision != "fp32" && a->precision != "bf16x3" && a->precision != "fp16" &&
      a->precision != "fp16x3" && a->precision != "plain") {
    FXH_LOG_ERROR("argument_invalid", "option=\"--precision\" value=\"%s\"", a->precision.c_str());
    return false;
  }
  if (a->debug) a->lazy_depth = 0;
  if (a->lazy_depth < 0) a->lazy_depth = 0;
  return true;
}

int precision_code(const std::string &p) {
  if (p == "bf16x3") return FXC_PREC_BF16X3;
  if (p == "fp16") return FXC_PREC_FP16;
  if (p == "fp16x3") return FXC_PREC_FP16X3;
  if (p == "plain") return FXC_PREC_PLAIN;
  return FXC_PREC_FP32;
}

// finalize_shape (native/xc/io/input.cu:1033-1054): nfft = 2(nspec-1),
// H = llround(C/dt) clamped to nspec-1, cc = 2H+1 <= nfft.
bool finalize_shape(int nspec, int nstep, float dt, float cclength, int *half_cc, int *cc) {
  if (nspec < 2 || nstep <= 0 || dt <= 0.f || cclength <= 0.f) return false;
  const int nfft = 2 * (nspec - 1);
  long long h = llround((double)cclength / (double)dt);
  if (h >= nspec) h = nspec - 1;
  const long long c = 2 * h + 1;
  if (c <= 0 || c > nfft) return false;
  *half_cc = (int)h;
  *cc = (int)c;
  return true;
}

// ------------------------------------------------------------ pack writing --
struct PackWriter {
  std::string root, ts, pack_path, tsv_path;
  size_t worker = 0, part = 0;
  XcJob job;
  uint64_t bytes = 0;
  FILE *pf = nullptr, *tf = nullptr;

  bool open_part() {
    char leaf[512];
    snprintf(leaf, sizeof(leaf), "%s.i%06zu.j%06zu.w%03zu.p%03zu", ts.c_str(), job.anchor_block,
             job.target_begin_block, worker, part);
    pack_path = root + "/" + leaf + ".xcpack";
    tsv_path = root + "/" + leaf + ".tsv";
    pf = fopen(pack_path.c_str(), "ab");
    tf = pf ? fopen(tsv_path.c_str(), "ab") : nullptr;
    if (!pf || !tf) {
      FXH_LOG_ERROR("xcpack_open_failed", "path=\"%s\" error=\"%s\"", pack_path.c_str(),
                    strerror(errno));
      return false;
    }
    fseeko(pf, 0, SEEK_END);
    fseeko(tf, 0, SEEK_END);
    if (ftello(tf) == 0)
      fputs("timestamp\tworker_id\tanchor_block\ttarget_begin_block\ttarget_end_block\t"
            "block_size\tanchor_begin\tanchor_end\ttarget_begin\ttarget_end\tpack_path\t"
            "offset\tbytes\tsrc_id\trec_id\tsrc_network\tsrc_station\tsrc_location\t"
            "src_component\trec_network\trec_station\trec_location\trec_component\tnpts\tdt\t"
            "dist\taz\tbaz\tfinal_pair_path\n",
            tf);
    bytes = (uint64_t)ftello(pf);
    return true;
  }
  bool close_part() {
    bool ok = true;
    if (pf) ok = fclose(pf) == 0 && ok;
    if (tf) ok = fclose(tf) == 0 && ok;
    pf = tf = nullptr;
    return ok;
  }
  bool append(const XcTask &t, const SpecMeta &src, const SpecMeta &rec, int npts, float dt,
              const std::string &final_path, const void *record, uint64_t nbytes) {
    if (bytes > 0 && bytes + nbytes > kMaxPackBytes) {
      if (!close_part()) return false;
      ++part;
      if (!open_part()) return false;
    }
    const uint64_t offset = bytes;
    if (fwrite(record, nbytes, 1, pf) != 1) return false;
    bytes += nbytes;
    fprintf(tf,
            "%s\t%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%zu\t%s\t%" PRIu64 "\t%" PRIu64
            "\t%d\t%d\t%s\t%s\t%s\t%s\t%s\t%s\t%s\t%s\t%d\t%.9g\t%.9g\t%.9g\t%.9g\t%s\n",
            clean_tsv_field(ts).c_str(), worker, job.anchor_block, job.target_begin_block,
            job.target_end_block, job.block_size, job.anchor_begin, job.anchor_end,
            job.target_begin, job.target_end, clean_tsv_field(pack_path).c_str(), offset, nbytes,
            (int)src.gnsl_id, (int)rec.gnsl_id, clean_tsv_field(src.network).c_str(),
            clean_tsv_field(src.station).c_str(), clean_tsv_field(src.location).c_str(),
            clean_tsv_field(src.component).c_str(), clean_tsv_field(rec.network).c_str(),
            clean_tsv_field(rec.station).c_str(), clean_tsv_field(rec.location).c_str(),
            clean_tsv_field(rec.component).c_str(), npts, (double)dt,
            (double)t.path.distance_km, (double)t.path.azimuth_deg,
            (double)t.path.back_azimuth_deg, clean_tsv_field(final_path).c_str());
    return true;
  }
};

// <out>/<sN>.<sS>/<rN>.<rS>/<sN>-<rN>.<sS>-<rS>.<sC>-<rC>.sac
// (native/xc/io/sac_record.cu:158-186; the directories are not created).
std::string final_pair_path(const std::string &out, const SpecMeta &s, const SpecMeta &r) {
  return out + "/" + s.network + "." + s.station + "/" + r.network + "." + r.station + "/" +
         s.network + "-" + r.network + "." + s.station + "-" + r.station + "." + s.component +
         "-" + r.component + ".sac";
}

// SacheadProcess + build_sac_header_for_task (native/xc/io/sac_header.c:3-38,
// native/xc/io/sac_record.cu:188-213).
SacHeader task_header(const XcTask &t, const SpecMeta &src, const SpecMeta &rec, float dt,
                      int cc, int half_cc, const TimeData &ti) {
  SacHeader h = sac_null();
  h.stla = rec.stla;
  h.stlo = rec.stlo;
  h.evla = src.stla;
  h.evlo = src.stlo;
  h.gcarc = t.path.great_circle_deg;
  h.az = t.path.azimuth_deg;
  h.baz = t.path.back_azimuth_deg;
  h.dist = t.path.distance_km;
  h.iftype = 1;
  h.leven = 1;
  h.delta = dt;
  h.npts = cc;
  const float cclen = (float)half_cc * dt;
  h.b = -1.0f * cclen;
  h.e = cclen;
  h.unused27 = 1;
  h.o = 0.f;
  h.nzyear = ti.year;
  h.nzjday = ti.day_of_year;
  h.nzhour = ti.hour;
  h.nzmin = ti.minute;
  h.nzsec = 0;
  h.nzmsec = 0;
  sac_text(h.kstnm, sizeof(h.kstnm), rec.station);
  sac_text(h.knetwk, sizeof(h.knetwk), rec.network);
  sac_text(h.kcmpnm, sizeof(h.kcmpnm), src.component + "-" + rec.component);
  sac_text(h.kevnm, sizeof(h.kevnm), src.network + "." + src.station);
  h.user0 = (float)src.gnsl_id;
  h.user1 = (float)rec.gnsl_id;
  h.user2 = (float)t.path_id;
  return h;
}

// One job's results waiting for the writer.
struct WriteBatch {
  std::shared_ptr<const TimestampWork> work;
  XcJob job;
  std::vector<XcTask> tasks;
  float *cc = nullptr;  // pinned, tasks x cc
  int slot = 0;
  int cc_size = 0, half_cc = 0;
  float dt = 0.f;
};

class Writer {
 public:
  Writer(const std::string &out, Progress *prog, int nslots) : out_(out), prog_(prog), free_(nslots) {
    for (int i = 0; i < nslots; ++i) free_[i] = i;
  }
  void start() { th_ = std::thread([this] { loop(); }); }
  // blocks until a pinned slot is free
  int acquire() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !free_.empty(); });
    const int s = free_.back();
    free_.pop_back();
    return s;
  }
  void submit(WriteBatch &&b, bool sync) {
    if (sync) {
      write(b);
      release(b.slot);
      return;
    }
    std::lock_guard<std::mutex> lk(mu_);
    q_.push_back(std::move(b));
    cv_.notify_all();
  }
  void drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return q_.empty() && !busy_; });
  }
  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      closed_ = true;
      cv_.notify_all();
    }
    if (th_.joinable()) th_.join();
  }
  bool failed() const { return failed_; }

 private:
  void release(int slot) {
    std::lock_guard<std::mutex> lk(mu_);
    free_.push_back(slot);
    cv_.notify_all();
  }
  void loop() {
    for (;;) {
      WriteBatch b;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return closed_ || !q_.empty(); });
        if (q_.empty()) return;
        b = std::move(q_.front());
        q_.pop_front();
        busy_ = true;
      }
      write(b);
      {
        std::lock_guard<std::mutex> lk(mu_);
        busy_ = false;
        free_.push_back(b.slot);
        cv_.notify_all();
      }
    }
  }
  void write(const WriteBatch &b) {
    if (b.tasks.empty()) return;
    TimeData ti;
    if (!parse_timestamp_text(b.work->timestamp, &ti))
      FXH_LOG_WARN("timestamp_parse_failed", "timestamp=\"%s\"", b.work->timestamp.c_str());
    PackWriter pw;
    pw.root = out_ + "/xcpack";
    pw.ts = b.work->timestamp;
    pw.job = b.job;
    if (!pw.open_part()) {
      failed_ = true;
      return;
    }
    std::vector<char> rec(sizeof(SacHeader) + (size_t)b.cc_size * sizeof(float));
    for (size_t k = 0; k < b.tasks.size(); ++k) {
      const XcTask &t = b.tasks[k];
      const SpecMeta &src = b.work->specs[t.src_meta];
      const SpecMeta &rcv = b.work->specs[t.rec_meta];
      const SacHeader h = task_header(t, src, rcv, b.dt, b.cc_size, b.half_cc, ti);
      memcpy(rec.data(), &h, sizeof(h));
      memcpy(rec.data() + sizeof(h), b.cc + k * (size_t)b.cc_size, (size_t)b.cc_size * 4);
      if (!pw.append(t, src, rcv, b.cc_size, b.dt, final_pair_path(out_, src, rcv), rec.data(),
                     rec.size())) {
        FXH_LOG_ERROR("xcpack_append_failed", "timestamp=\"%s\" path_id=%d",
                      b.work->timestamp.c_str(), t.path_id);
        failed_ = true;
        break;
      }
    }
    if (!pw.close_part()) failed_ = true;
    prog_->add("current", b.tasks.size(), b.work->timestamp);
  }

  std::string out_;
  Progress *prog_;
  std::vector<int> free_;
  std::deque<WriteBatch> q_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::thread th_;
  bool closed_ = false, busy_ = false, failed_ = false;
};

#define HIPX(call)                                                                   \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      FXH_LOG_ERROR("hip_error", "call=\"%s\" error=\"%s\"", #call, hipGetErrorString(e_)); \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

#define FXCX(call)                                                                      \
  do {                                                                                  \
    if ((call) != FXC_OK) {                                                             \
      FXH_LOG_ERROR("fxc_error", "call=\"%s\" error=\"%s\"", #call, fxc_last_error()); \
      return 1;                                                                         \
    }                                                                                   \
  } while (0)

size_t block_size_for(const Args &a, size_t rows) {
  size_t b = std::min(rows, kMaxBlockFiles);
  if (a.block_files > 0) b = std::min(b, a.block_files);
  return std::max<size_t>(1, b);
}

int dry_run(const Args &a, const std::vector<TimestampInput> &inputs,
            const std::vector<PathRecord> &table) {
  printf("timestamp\tanchor_block\ttarget_begin_block\tsrc_row\trec_row\tsrc_local\trec_local\tpath_id\tautocorr\n");
  for (const auto &in : inputs) {
    TimestampWork w;
    if (!open_timestamp_work(in, &w)) return 1;
    int half = 0, cc = 0;
    if (!finalize_shape(w.nspec, w.nstep, w.dt, a.cclength, &half, &cc)) return 1;
    const size_t block = block_size_for(a, w.specs.size());
    for (const XcJob &j : build_jobs(w.specs.size(), block)) {
      std::vector<XcTask> tasks;
      std::vector<size_t> loaded;
      build_tasks_for_job(w.specs, table, j, &tasks, &loaded);
      for (const auto &t : tasks)
        printf("%s\t%zu\t%zu\t%zu\t%zu\t%d\t%d\t%d\t%d\n", w.timestamp.c_str(), j.anchor_block,
               j.target_begin_block, t.src_meta, t.rec_meta, t.src_local, t.rec_local, t.path_id,
               t.is_autocorr ? 1 : 0);
    }
  }
  return 0;
}

int run(const Args &a) {
  std::vector<PathRecord> table;
  if (!load_path_table(a.paths, &table)) return 1;
  std::vector<TimestampInput> inputs;
  if (!load_timestamp_inputs(a.input, &inputs)) return 1;
  if (a.dry_run) return dry_run(a, inputs, table);

  const std::string pack_root = a.out + "/xcpack";
  if (mkdir_p(pack_root) != 0) {
    FXH_LOG_ERROR("xcpack_dir_create_failed", "path=\"%s\"", pack_root.c_str());
    return 1;
  }
  // indexes of finished timestamps are built on a side thread while the
  // next timestamp correlates (the reference polls the .done markers)
  std::unique_ptr<SourcePackMaterializer> sp;
  if (!a.sourcepack_dir.empty()) {
    sp.reset(new SourcePackMaterializer(a.out, a.sourcepack_dir, a.sourcepack_sort,
                                        a.sourcepack_progress));
    sp->start();
  }
  Progress prog;
  prog.init(a.progress.empty() ? nullptr : a.progress.c_str());
  prog.set_rows({{"overall", "RUNNING", 0, inputs.size(), "timestamps", ""},
                 {"current", "PENDING", 0, 0, "xc_units", ""}});

  const int dev = atoi(a.gpu_ids.c_str());
  HIPX(hipSetDevice(dev));
  hipStream_t st;
  HIPX(hipStreamCreate(&st));
  size_t mem_free = 0, mem_total = 0;
  HIPX(hipMemGetInfo(&mem_free, &mem_total));
  size_t budget = (size_t)(0.90 * (double)mem_free);
  const long long cap_mib = a.mem_caps.empty() ? 0 : atoll(a.mem_caps.c_str());
  if (cap_mib > 0) budget = std::min(budget, (size_t)cap_mib << 20);

  const int nslots = a.lazy_depth + 1;
  Writer writer(a.out, &prog, nslots);
  if (a.lazy_depth > 0) writer.start();
  std::vector<float *> cc_host(nslots, nullptr);  // pinned correlation buffers, one per slot
  std::vector<size_t> slot_cap(nslots, 0);
  std::complex<float> *payload = nullptr, *gather = nullptr;
  size_t payload_elems = 0;
  float2 *d_spec = nullptr;
  float *d_cc = nullptr;
  size_t d_spec_elems = 0, d_cc_elems = 0;
  fxc_xc_plan *plan = nullptr;
  int plan_key[4] = {0, 0, 0, 0};
  float plan_dt = 0.f;
  int rc = 0;

  auto cleanup = [&]() {
    if (a.lazy_depth > 0) {
      writer.drain();
      writer.stop();
    }
    if (plan) fxc_xc_plan_destroy(plan);
    for (float *p : cc_host)
      if (p) (void)hipHostFree(p);
    if (payload) (void)hipHostFree(payload);
    if (gather) (void)hipHostFree(gather);
    if (d_spec) (void)hipFree(d_spec);
    if (d_cc) (void)hipFree(d_cc);
    (void)hipStreamDestroy(st);
  };

  for (size_t ti = 0; ti < inputs.size() && rc == 0; ++ti) {
    auto work = std::make_shared<TimestampWork>();
    if (!open_timestamp_work(inputs[ti], work.get())) {
      FXH_LOG_WARN("timestamp_open_failed", "timestamp=\"%s\"", inputs[ti].timestamp.c_str());
      prog.update("current", "SKIPPED", 0, 0, "xc_units", inputs[ti].timestamp);
      prog.update("overall", ti + 1 == inputs.size() ? "DONE" : "RUNNING", ti + 1, inputs.size(),
                  "timestamps", inputs[ti].timestamp);
      continue;
    }
    int half = 0, cc = 0;
    if (!finalize_shape(work->nspec, work->nstep, work->dt, a.cclength, &half, &cc)) {
      FXH_LOG_ERROR("shape_invalid", "timestamp=\"%s\" nspec=%d nstep=%d dt=%.9g",
                    work->timestamp.c_str(), work->nspec, work->nstep, (double)work->dt);
      rc = 1;
      break;
    }
    const size_t rows = work->specs.size();
    const size_t need = (size_t)work->nstep * rows * work->nspec;
    if (need > payload_elems) {
      if (payload) (void)hipHostFree(payload);
      if (gather) (void)hipHostFree(gather);
      payload = gather = nullptr;
      if (hipHostMalloc((void **)&payload, need * 8) != hipSuccess ||
          hipHostMalloc((void **)&gather, need * 8) != hipSuccess) {
        FXH_LOG_ERROR("host_alloc_failed", "bytes=%zu", need * 8);
        rc = 1;
        break;
      }
      payload_elems = need;
    }
    if (!read_timestamp_payload(*work, payload)) {
      rc = 1;
      break;
    }
    // plan per shape
    const int prec = precision_code(a.precision);
    if (!plan || plan_key[0] != work->nspec || plan_key[1] != work->nstep || plan_key[2] != half ||
        plan_key[3] != prec || plan_dt != work->dt) {
      if (plan) fxc_xc_plan_destroy(plan);
      plan = nullptr;
      FXCX(fxc_xc_plan_create(&plan, work->nspec, work->nstep, work->dt, half, prec));
      plan_key[0] = work->nspec;
      plan_key[1] = work->nstep;
      plan_key[2] = half;
      plan_key[3] = prec;
      plan_dt = work->dt;
    }
    const size_t block = block_size_for(a, rows);
    const std::vector<XcJob> jobs = build_jobs(rows, block);
    // units of this timestamp: all tasks of all jobs
    std::vector<std::vector<XcTask>> job_tasks(jobs.size());
    std::vector<std::vector<size_t>> job_loaded(jobs.size());
    size_t units = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
      build_tasks_for_job(work->specs, table, jobs[j], &job_tasks[j], &job_loaded[j]);
      units += job_tasks[j].size();
    }
    prog.update("current", units == 0 ? "DONE" : "RUNNING", 0, units, "xc_units", work->timestamp);
    FXH_LOG_INFO("timestamp_plan", "timestamp=\"%s\" files=%zu block_files=%zu row_jobs=%zu units=%zu",
                 work->timestamp.c_str(), rows, block, jobs.size(), units);
    for (size_t j = 0; j < jobs.size() && rc == 0; ++j) {
      const auto &tasks = job_tasks[j];
      if (tasks.empty()) continue;
      const auto &loaded = job_loaded[j];
      const size_t nl = loaded.size();
      // spectra of the loaded rows, [nstep][nl][nspec]
      const std::complex<float> *src = payload;
      if (!(nl == rows && loaded.front() == 0)) {
        for (int s = 0; s < work->nstep; ++s)
          for (size_t r = 0; r < nl; ++r)
            memcpy(gather + ((size_t)s * nl + r) * work->nspec,
                   payload + ((size_t)s * rows + loaded[r]) * work->nspec, (size_t)work->nspec * 8);
        src = gather;
      }
      const size_t spec_elems = (size_t)work->nstep * nl * work->nspec;
      const size_t cc_elems = tasks.size() * (size_t)cc;
      if (spec_elems > d_spec_elems) {
        if (d_spec) (void)hipFree(d_spec);
        d_spec = nullptr;
        HIPX(hipMalloc((void **)&d_spec, spec_elems * 8));
        d_spec_elems = spec_elems;
      }
      if (cc_elems > d_cc_elems) {
        if (d_cc) (void)hipFree(d_cc);
        d_cc = nullptr;
        HIPX(hipMalloc((void **)&d_cc, cc_elems * 4));
        d_cc_elems = cc_elems;
      }
      // workspace rows that fit the remaining budget (the plan chunks the rest)
      const size_t used = d_spec_elems * 8 + d_cc_elems * 4 +
                          (prec == FXC_PREC_BF16X3 || prec == FXC_PREC_FP16 ? nl * (size_t)work->nspec * 256 : 0);
      const size_t pitch = ((size_t)work->nspec + 15) / 16 * 16 * 8;
      long long chunk = 0;
      if (budget > used + (256ull << 20)) chunk = (long long)((budget - used - (256ull << 20)) / pitch);
      if (chunk < 1024) chunk = 1024;
      if ((size_t)chunk >= tasks.size()) chunk = 0;
      FXCX(fxc_xc_set_chunk_rows(plan, chunk));
      std::vector<int> s_idx(tasks.size()), r_idx(tasks.size());
      for (size_t k = 0; k < tasks.size(); ++k) {
        s_idx[k] = tasks[k].src_local;
        r_idx[k] = tasks[k].rec_local;
      }
      FXCX(fxc_xc_set_pairs(plan, (int)nl, s_idx.data(), r_idx.data(), (long long)tasks.size()));
      HIPX(hipMemcpyAsync(d_spec, src, spec_elems * 8, hipMemcpyHostToDevice, st));
      FXCX(fxc_xc_correlate(plan, (const float *)d_spec, d_cc, st));
      const int slot = writer.acquire();
      float *&hb = cc_host[slot];
      if (slot_cap[slot] < cc_elems) {
        if (hb) (void)hipHostFree(hb);
        hb = nullptr;
        HIPX(hipHostMalloc((void **)&hb, cc_elems * 4));
        slot_cap[slot] = cc_elems;
      }
      HIPX(hipMemcpyAsync(hb, d_cc, cc_elems * 4, hipMemcpyDeviceToHost, st));
      HIPX(hipStreamSynchronize(st));
      WriteBatch wb;
      wb.work = work;
      wb.job = jobs[j];
      wb.tasks = tasks;
      wb.cc = hb;
      wb.slot = slot;
      wb.cc_size = cc;
      wb.half_cc = half;
      wb.dt = work->dt;
      writer.submit(std::move(wb), a.lazy_depth == 0);
    }
    if (rc) break;
    if (a.lazy_depth > 0) writer.drain();
    if (writer.failed()) {
      rc = 1;
      break;
    }
    if (write_marker(pack_root + "/" + work->timestamp + ".done") != 0) {
      FXH_LOG_ERROR("xcpack_timestamp_done_failed", "timestamp=\"%s\" output=\"%s\"",
                    work->timestamp.c_str(), a.out.c_str());
      rc = 1;
      break;
    }
    if (sp) sp->submit(work->timestamp);
    prog.update("current", "DONE", units, units, "xc_units", work->timestamp);
    prog.update("overall", ti + 1 == inputs.size() ? "DONE" : "RUNNING", ti + 1, inputs.size(),
                "timestamps", work->timestamp);
    FXH_LOG_INFO("timestamp_done", "timestamp_index=%zu timestamp=\"%s\"", ti,
                 work->timestamp.c_str());
  }
  cleanup();
  if (rc == 0 && write_marker(pack_root + "/_SUCCESS") != 0) {
    FXH_LOG_ERROR("xcpack_success_marker_failed", "output=\"%s\"", a.out.c_str());
    rc = 1;
  }
  if (sp && !sp->finish(rc == 0, "upstream XC did not complete; global success marker withheld"))
    rc = 1;
  prog.finish(rc == 0 ? "DONE" : "FAILED", rc == 0);
  FXH_LOG_INFO("run_done", "timestamps=%zu", inputs.size());
  return rc;
}

}  // namespace

int main(int argc, char **argv) {
  Args a;
  int code = 1;
  if (!parse_args(argc, argv, &a, &code)) return code;
  FXH_LOG_INFO("run_start", "output=\"%s\" progress=\"%s\" precision=%s output_mode=xcpack",
               a.out.c_str(), a.progress.c_str(), a.precision.c_str());
  return run(a);
}
