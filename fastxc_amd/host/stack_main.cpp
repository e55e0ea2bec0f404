// ncf_pws and ncf_tfpws: phase-weighted stacks of SourcePack records on one
// MI355X (this file builds both; FXH_TFPWS selects the Stockwell form).
//
// Command lines, inputs and outputs follow the reference executables
// (native/pws/arguproc.c:5-105, native/pws/worker/sourcepack_pipeline.cu:340-470,
// native/pws/io/sourcepack_io.cu; native/tfpws/arguproc.c:36-103,
// native/tfpws/worker/pipeline.cu:200-327, native/tfpws/io/sourcepack_io.cu):
//   ncf_pws   --sourcepack-list f --output-sourcepack dir [--progress f]
//             [-G ids] [-M MiB,...] [-B substack] [-H staged_group_limit]
//   ncf_tfpws -I f -O dir [--progress f] [-G ids] [-M MiB] [-S 001] [-B n]
//             [-F fmin/fmax [-T taper_hz]]
// The list file names sourcepack_index.tsv files; their rows (sorted by
// path_id, component_slot) are merged, each (path_id, component_slot) group
// is one pair, its records sorted by timestamp form sub-stacks of B records.
// Outputs: dir/{pws,tfpws}.w000.{pack,tsv}, dir/sourcepack_index.tsv (25
// columns, timestamp "STACK"), dir/_SUCCESS, and the progress sidecar.
#include "fxc_host.hpp"

#include "fastxc_amd.h"

#include <hip/hip_runtime.h>

#include <getopt.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

using namespace fxh;

namespace {

#ifdef FXH_TFPWS
constexpr bool kTf = true;
const char *const kName = "tfpws";
#else
constexpr bool kTf = false;
const char *const kName = "pws";
#endif

const char *const kIndexHeader =
    "timestamp\tpath_id\tcomponent_slot\tsource_key\treceiver_key\tsrc_id\trec_id\t"
    "src_network\tsrc_station\tsrc_location\tsrc_component\trec_network\trec_station\t"
    "rec_location\trec_component\tnpts\tdt\tdist\taz\tbaz\trecord_path\trecord_offset\tbytes\t"
    "storage_kind\tfinal_pair_path";

struct Args {
  std::string list, out, progress, gpu = "0", mem;
  int substack = 1;
  size_t staged_limit = 200000;
  bool band = false;
  double fmin = 0, fmax = 0, taper = 0;
  float power = 1.f;  // spec mode: exponent of |W| (the reference: 1)
  int voices = 0;     // spec mode (TF-PWS): voice count (0: every voice, the reference)
};

void usage() {
  if (kTf)
    fprintf(stderr,
            "Usage: ncf_tfpws -I sourcepack_list -O output_sourcepack [options]\n"
            "  -I, --sourcepack-list   file listing sourcepack_index.tsv paths\n"
            "  -O, --output-sourcepack output directory\n"
            "  -G  GPU list (the first id is used)   -M  MiB cap\n"
            "  -S  output selection, must end in 1 (e.g. 001)\n"
            "  -B  sub-stack size (default 1)\n"
            "  -F  fmin/fmax band for the band-limited blend   -T taper (Hz)\n"
            "  --progress <file>\n"
            "  --power <p>   spec mode: |W|^p (default 1, the reference)\n"
            "  --voices <v>  spec mode: v voices, weights interpolated between them\n"
            "                (default 0: every voice, the reference)\n");
  else
    fprintf(stderr,
            "Usage: ncf_pws --sourcepack-list f --output-sourcepack dir [options]\n"
            "  -G, --gpu-workers <ids>        GPU list (the first id is used)\n"
            "  -M, --gpu-memory-mib <MiB>     device memory cap\n"
            "  -B, --substack-size <n>        records per sub-stack (default 1)\n"
            "  -H, --staged-group-limit <n>   host staging bound in groups\n"
            "  --progress <file>\n"
            "  --power <p>                    spec mode: |W|^p (default 1, the reference)\n");
}

bool parse_args(int argc, char **argv, Args *a) {
  if (argc <= 1) {
    usage();
    return false;
  }
  static struct option longs[] = {{"progress", required_argument, 0, 1000},
                                  {"progress-file", required_argument, 0, 1000},
                                  {"index-list", required_argument, 0, 1001},
                                  {"sourcepack-list", required_argument, 0, 1001},
                                  {"output-dir", required_argument, 0, 1002},
                                  {"output-sourcepack", required_argument, 0, 1002},
                                  {"gpu-workers", required_argument, 0, 'G'},
                                  {"gpu-memory-mib", required_argument, 0, 'M'},
                                  {"substack-size", required_argument, 0, 'B'},
                                  {"staged-group-limit", required_argument, 0, 'H'},
                                  {"host-group-limit", required_argument, 0, 'H'},
                                  {"power", required_argument, 0, 1003},
                                  {"voices", required_argument, 0, 1004},
                                  {0, 0, 0, 0}};
  const char *shorts = kTf ? "I:O:G:S:B:M:F:T:" : "G:B:H:M:";
  int c, idx = 0;
  opterr = 0;
  while ((c = getopt_long(argc, argv, shorts, longs, &idx)) != -1) {
    switch (c) {
      case 'I':
      case 1001: a->list = optarg; break;
      case 'O':
      case 1002: a->out = optarg; break;
      case 1000: a->progress = optarg; break;
      case 'G':
        a->gpu = optarg;
        if (a->gpu.empty()) {
          FXH_LOG_ERROR("invalid_gpu_workers", "message=\"-G must not be empty\"");
          return false;
        }
        break;
      case 'M': a->mem = optarg; break;
      case 1003: {
        char *end = nullptr;
        a->power = strtof(optarg, &end);
        if (!end || *end || !(a->power > 0.f) || !std::isfinite(a->power)) {
          FXH_LOG_ERROR("invalid_power", "value=\"%s\"", optarg);
          return false;
        }
        break;
      }
      case 1004:
        a->voices = atoi(optarg);
        if (!kTf || a->voices < 0 || a->voices == 1) {
          FXH_LOG_ERROR("invalid_voices", "value=\"%s\" (ncf_tfpws only; 0 or >= 2)", optarg);
          return false;
        }
        break;
      case 'B':
        a->substack = atoi(optarg);
        if (a->substack < 2) a->substack = 1;
        break;
      case 'H':
        a->staged_limit = (size_t)strtoull(optarg, nullptr, 10);
        if (a->staged_limit < 1) {
          FXH_LOG_ERROR("invalid_staged_group_limit", "value=%zu", a->staged_limit);
          return false;
        }
        break;
      case 'S':
        if (strlen(optarg) != 3 || strspn(optarg, "01") != 3 || optarg[2] != '1') {
          FXH_LOG_ERROR("invalid_output_selection", "value=\"%s\" expected=\"??1\"", optarg);
          return false;
        }
        if (optarg[0] == '1' || optarg[1] == '1')
          FXH_LOG_WARN("output_selection_reduced", "value=\"%s\" written=tfpws", optarg);
        break;
      case 'F':
        if (sscanf(optarg, "%lf/%lf", &a->fmin, &a->fmax) != 2 || a->fmin < 0.0 ||
            a->fmax <= a->fmin) {
          FXH_LOG_ERROR("invalid_band", "value=\"%s\"", optarg);
          return false;
        }
        a->band = true;
        break;
      case 'T':
        a->taper = atof(optarg);
        if (a->taper < 0.0) {
          FXH_LOG_ERROR("invalid_band_taper", "value=\"%s\"", optarg);
          return false;
        }
        break;
      default:
        FXH_LOG_ERROR("unknown_option", "option=\"%s\"", argv[optind - 1]);
        usage();
        return false;
    }
  }
  if (a->list.empty() || a->out.empty()) {
    FXH_LOG_ERROR("missing_sourcepack_io", "index_list=\"%s\" output_dir=\"%s\"", a->list.c_str(),
                  a->out.c_str());
    usage();
    return false;
  }
  return true;
}

// --------------------------------------------------------- SourcePack rows --
struct Record {
  std::map<std::string, std::string> f;  // field -> value (trimmed)
  std::string path_id;
  int slot = 0;
  std::string record_path;
  long long offset = 0, bytes = 0;
  const std::string &get(const char *k) const {
    static const std::string empty;
    auto it = f.find(k);
    return it == f.end() ? empty : it->second;
  }
};

class Stream {
 public:
  explicit Stream(const std::string &p) : path_(p) {}
  bool open() {
    in_.open(path_);
    std::string header;
    if (!in_ || !std::getline(in_, header)) {
      FXH_LOG_ERROR("open_sourcepack_index_failed", "path=\"%s\"", path_.c_str());
      return false;
    }
    names_ = split_tab(header);
    for (auto &n : names_) n = trim(n);
    static const char *const req[] = {
        "timestamp", "path_id", "component_slot", "source_key", "receiver_key", "src_id",
        "rec_id", "src_network", "src_station", "src_location", "src_component", "rec_network",
        "rec_station", "rec_location", "rec_component", "npts", "dt", "dist", "az", "baz",
        "record_path", "record_offset", "bytes", "final_pair_path"};
    for (const char *r : req)
      if (column_index(names_, r) < 0) {
        FXH_LOG_ERROR("sourcepack_index_missing_field", "path=\"%s\" field=%s", path_.c_str(), r);
        return false;
      }
    return true;
  }
  bool next(Record *r) {
    std::string line;
    while (std::getline(in_, line)) {
      if (line.empty()) continue;
      auto row = split_tab(line);
      r->f.clear();
      for (size_t i = 0; i < names_.size(); ++i) r->f[names_[i]] = i < row.size() ? trim(row[i]) : "";
      r->path_id = r->get("path_id");
      r->slot = atoi(r->get("component_slot").c_str());
      r->record_path = r->get("record_path");
      r->offset = strtoll(r->get("record_offset").c_str(), nullptr, 10);
      r->bytes = strtoll(r->get("bytes").c_str(), nullptr, 10);
      return true;
    }
    return false;
  }

 private:
  std::string path_;
  std::ifstream in_;
  std::vector<std::string> names_;
};

// k-way merge by (path_id text, component_slot, arrival order); one group =
// equal (path_id, slot), its records sorted by (timestamp, path, offset).
class GroupReader {
 public:
  explicit GroupReader(const std::vector<std::string> &paths) {
    for (const auto &p : paths) streams_.emplace_back(new Stream(p));
  }
  bool open() {
    for (size_t i = 0; i < streams_.size(); ++i) {
      if (!streams_[i]->open()) return false;
      Record r;
      if (streams_[i]->next(&r)) push(r, i);
    }
    return true;
  }
  bool next_group(std::vector<Record> *g) {
    g->clear();
    if (heap_.empty()) return false;
    const std::string pid = heap_.top().rec.path_id;
    const int slot = heap_.top().rec.slot;
    while (!heap_.empty() && heap_.top().rec.path_id == pid && heap_.top().rec.slot == slot) {
      Item it = heap_.top();
      heap_.pop();
      g->push_back(it.rec);
      Record nx;
      if (streams_[it.stream]->next(&nx)) push(nx, it.stream);
    }
    std::sort(g->begin(), g->end(), [](const Record &a, const Record &b) {
      if (a.get("timestamp") != b.get("timestamp")) return a.get("timestamp") < b.get("timestamp");
      if (a.record_path != b.record_path) return a.record_path < b.record_path;
      return a.offset < b.offset;
    });
    return true;
  }

 private:
  struct Item {
    Record rec;
    size_t stream, seq;
  };
  struct Cmp {
    bool operator()(const Item &a, const Item &b) const {
      if (a.rec.path_id != b.rec.path_id) return a.rec.path_id > b.rec.path_id;
      if (a.rec.slot != b.rec.slot) return a.rec.slot > b.rec.slot;
      return a.seq > b.seq;
    }
  };
  void push(const Record &r, size_t s) { heap_.push(Item{r, s, seq_++}); }
  std::vector<std::unique_ptr<Stream>> streams_;
  std::priority_queue<Item, std::vector<Item>, Cmp> heap_;
  size_t seq_ = 0;
};

bool read_list(const std::string &path, std::vector<std::string> *out) {
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    out->push_back(line);
  }
  if (out->empty()) FXH_LOG_ERROR("sourcepack_list_empty", "path=\"%s\"", path.c_str());
  return !out->empty();
}

// Samples of a record's SAC trace (IXY traces carry x and y).
unsigned data_count(const SacHeader &h) { return h.iftype == 4 ? (unsigned)(h.npts * 2) : (unsigned)h.npts; }

bool read_record(const Record &r, SacHeader *h, std::vector<float> *data, unsigned expect) {
  FILE *fp = fopen(r.record_path.c_str(), "rb");
  bool ok = fp && fseeko(fp, (off_t)r.offset, SEEK_SET) == 0 && fread(h, sizeof(*h), 1, fp) == 1;
  if (ok) {
    const unsigned n = data_count(*h);
    ok = (expect == 0 || n == expect) &&
         r.bytes == (long long)sizeof(SacHeader) + (long long)n * 4;
    if (ok) {
      data->resize(n);
      ok = fread(data->data(), 4, n, fp) == n;
    }
  }
  if (fp) fclose(fp);
  if (!ok)
    FXH_LOG_ERROR("sourcepack_record_invalid", "path=\"%s\" offset=%lld", r.record_path.c_str(),
                  r.offset);
  return ok;
}

// One pair: sub-stacks of `substack` consecutive records (timestamp order).
// PWS (prepare_sourcepack_host_item, native/pws/io/sourcepack_io.cu:385-478):
// rows are group means, w_g = n_g/T, lin = sum_g w_g mean_g, all float.
// TF-PWS (native/tfpws/io/sourcepack_io.cu:403-490): rows are group sums,
// w_g = n_g/T, lin = sum_g sum_g/T.
struct Pair {
  Record rec;
  SacHeader hdr;
  unsigned n = 0, ngroups = 0, nseg = 0;
  std::vector<float> rows, weights, lin;
};

bool prepare_pair(const std::vector<Record> &recs, int substack, Pair *p) {
  p->rec = recs[0];
  std::vector<float> buf;
  if (!read_record(recs[0], &p->hdr, &buf, 0)) return false;
  p->n = data_count(p->hdr);
  const size_t gsz = substack < 2 ? 1 : (size_t)substack;
  p->nseg = (unsigned)recs.size();
  p->ngroups = (unsigned)((recs.size() + gsz - 1) / gsz);
  p->rows.assign((size_t)p->ngroups * p->n, 0.f);
  p->weights.assign(p->ngroups, 0.f);
  p->lin.assign(p->n, 0.f);
  const float inv_total = 1.0f / (float)p->nseg;
  for (unsigned g = 0; g < p->ngroups; ++g) {
    const size_t first = g * gsz, last = std::min(first + gsz, recs.size());
    float *tr = p->rows.data() + (size_t)g * p->n;
    for (size_t i = first; i < last; ++i) {
      SacHeader h;
      if (!read_record(recs[i], &h, &buf, p->n)) return false;
      for (unsigned j = 0; j < p->n; ++j) tr[j] += buf[j];
    }
    const size_t cnt = last - first;
    if (kTf) {
      p->weights[g] = (float)cnt * inv_total;
      for (unsigned j = 0; j < p->n; ++j) p->lin[j] += tr[j] * inv_total;
    } else {
      const float inv = 1.0f / (float)cnt;
      for (unsigned j = 0; j < p->n; ++j) tr[j] *= inv;
      p->weights[g] = (float)cnt / (float)p->nseg;
      for (unsigned j = 0; j < p->n; ++j) p->lin[j] += tr[j] * p->weights[g];
    }
  }
  if (!kTf) {
    // configure_output_header (native/pws/io/sac_output.hpp)
    SacHeader &h = p->hdr;
    h.npts = (int)p->n;
    h.iftype = 1;
    h.leven = 1;
    if (h.delta != -12345.0f) h.e = h.b + (float)(p->n - 1) * h.delta;
    h.user0 = (float)p->nseg;
    h.user1 = (float)p->ngroups;
    h.user2 = (float)gsz;
    sac_text(h.kuser0, 8, "NTRACE");
    sac_text(h.kuser1, 8, "NGROUP");
    sac_text(h.kuser2, 8, "SUBSTK");
    h.nzyear = 2010;
    h.nzjday = 214;
    h.nzhour = h.nzmin = h.nzsec = h.nzmsec = 0;
  }
  return true;
}

std::string final_pair_path(const Record &r) {
  std::string p = r.get("final_pair_path");
  if (p.empty())
    p = r.get("src_network") + "-" + r.get("rec_network") + "." + r.get("src_station") + "-" +
        r.get("rec_station") + "." + r.get("src_component") + "-" + r.get("rec_component");
  const std::string sfx = ".sac", repl = kTf ? ".tfpws.sac" : ".pws.sac";
  if (p.size() >= sfx.size() && p.compare(p.size() - sfx.size(), sfx.size(), sfx) == 0)
    p.replace(p.size() - sfx.size(), sfx.size(), repl);
  else
    p += repl;
  return p;
}

std::string fmt_stream(float v) {
  std::ostringstream s;
  s << v;
  return s.str();
}

#define HIPX(call)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      FXH_LOG_ERROR("hip_error", "call=\"%s\" error=\"%s\"", #call, hipGetErrorString(e_)); \
      return false;                                                                          \
    }                                                                                        \
  } while (0)

class Runner {
 public:
  Runner(const Args &a, const std::string &out) : a_(a), out_(out) {}
  bool init() {
    HIPX(hipSetDevice(atoi(a_.gpu.c_str())));
    HIPX(hipStreamCreate(&st_));
    size_t fr = 0, tot = 0;
    HIPX(hipMemGetInfo(&fr, &tot));
    budget_ = (size_t)(0.5 * (double)fr);
    const long long cap = a_.mem.empty() ? 0 : atoll(a_.mem.c_str());
    if (cap > 0) budget_ = std::min(budget_, (size_t)cap << 20);
    char name[64];
    snprintf(name, sizeof(name), "%s.w000.pack", kName);
    pack_path_ = path_join(out_, name);
    snprintf(name, sizeof(name), "%s.w000.tsv", kName);
    shard_path_ = path_join(out_, name);
    pack_ = fopen(pack_path_.c_str(), "wb");
    shard_.open(shard_path_);
    if (!pack_ || !shard_) {
      FXH_LOG_ERROR("shard_open_failed", "path=\"%s\"", pack_path_.c_str());
      return false;
    }
    shard_ << kIndexHeader << '\n';
    return true;
  }
  ~Runner() {
    if (pack_) fclose(pack_);
    for (void *p : {(void *)d_rows_, (void *)d_w_, (void *)d_lin_, (void *)d_out_})
      if (p) (void)hipFree(p);
    if (st_) (void)hipStreamDestroy(st_);
  }
  // queue a pair; runs the batch when it is full
  bool add(Pair &&p) {
    // one launch shares n (and, for the Stockwell band, delta)
    if (!batch_.empty() && (batch_[0].n != p.n || (kTf && batch_[0].hdr.delta != p.hdr.delta) ||
                            !fits(p))) {
      if (!flush()) return false;
    }
    groups_ += p.ngroups;
    batch_.push_back(std::move(p));
    if (groups_ >= a_.staged_limit) return flush();
    return true;
  }
  bool flush() {
    if (batch_.empty()) return true;
    const int n = (int)batch_[0].n, np = (int)batch_.size();
    std::vector<long long> offs(np + 1, 0);
    for (int i = 0; i < np; ++i) offs[i + 1] = offs[i] + batch_[i].ngroups;
    const size_t rows = (size_t)offs[np];
    std::vector<float> h_rows(rows * n), h_w(rows), h_lin((size_t)np * n), h_out((size_t)np * n);
    for (int i = 0; i < np; ++i) {
      std::copy(batch_[i].rows.begin(), batch_[i].rows.end(), h_rows.begin() + offs[i] * n);
      std::copy(batch_[i].weights.begin(), batch_[i].weights.end(), h_w.begin() + offs[i]);
      std::copy(batch_[i].lin.begin(), batch_[i].lin.end(), h_lin.begin() + (size_t)i * n);
    }
    if (!ensure(&d_rows_, &cap_rows_, rows * n) || !ensure(&d_w_, &cap_w_, rows) ||
        !ensure(&d_lin_, &cap_lin_, (size_t)np * n) || !ensure(&d_out_, &cap_out_, (size_t)np * n))
      return false;
    HIPX(hipMemcpyAsync(d_rows_, h_rows.data(), rows * n * 4, hipMemcpyHostToDevice, st_));
    HIPX(hipMemcpyAsync(d_w_, h_w.data(), rows * 4, hipMemcpyHostToDevice, st_));
    HIPX(hipMemcpyAsync(d_lin_, h_lin.data(), (size_t)np * n * 4, hipMemcpyHostToDevice, st_));
    int rc;
    const fxc_stack_spec spec{a_.power, a_.voices};
    const bool parity = a_.power == 1.f && a_.voices == 0;
    if (kTf)
      rc = parity ? fxc_tfpws(d_rows_, d_w_, offs.data(), d_lin_, np, n, a_.band ? 1 : 0,
                              batch_[0].hdr.delta, (float)a_.fmin, (float)a_.fmax,
                              (float)a_.taper, d_out_, st_)
                  : fxc_tfpws_spec(d_rows_, d_w_, offs.data(), d_lin_, np, n, a_.band ? 1 : 0,
                                   batch_[0].hdr.delta, (float)a_.fmin, (float)a_.fmax,
                                   (float)a_.taper, &spec, d_out_, st_);
    else
      rc = parity ? fxc_pws(d_rows_, d_w_, offs.data(), d_lin_, np, n, d_out_, st_)
                  : fxc_pws_spec(d_rows_, d_w_, offs.data(), d_lin_, np, n, &spec, d_out_, st_);
    if (rc != FXC_OK) {
      FXH_LOG_ERROR("stack_compute_failed", "error=\"%s\"", fxc_last_error());
      return false;
    }
    HIPX(hipMemcpyAsync(h_out.data(), d_out_, (size_t)np * n * 4, hipMemcpyDeviceToHost, st_));
    HIPX(hipStreamSynchronize(st_));
    for (int i = 0; i < np; ++i) {
      Pair &p = batch_[i];
      const float *d = h_out.data() + (size_t)i * n;
      if (!kTf) {
        // update_output_stats
        float mn = d[0], mx = d[0];
        double sum = 0;
        for (int j = 0; j < n; ++j) {
          mn = std::min(mn, d[j]);
          mx = std::max(mx, d[j]);
          sum += d[j];
        }
        p.hdr.depmin = mn;
        p.hdr.depmax = mx;
        p.hdr.depmen = (float)(sum / (double)n);
      }
      const long long off = (long long)ftello(pack_);
      if (fwrite(&p.hdr, sizeof(p.hdr), 1, pack_) != 1 || fwrite(d, 4, (size_t)n, pack_) != (size_t)n) {
        FXH_LOG_ERROR("write_pack_failed", "path=\"%s\"", pack_path_.c_str());
        return false;
      }
      const long long bytes = (long long)sizeof(SacHeader) + (long long)n * 4;
      const Record &r = p.rec;
      shard_ << "STACK" << '\t' << r.path_id << '\t' << r.slot << '\t' << r.get("source_key") << '\t'
             << r.get("receiver_key") << '\t' << r.get("src_id") << '\t' << r.get("rec_id") << '\t'
             << r.get("src_network") << '\t' << r.get("src_station") << '\t'
             << r.get("src_location") << '\t' << r.get("src_component") << '\t'
             << r.get("rec_network") << '\t' << r.get("rec_station") << '\t'
             << r.get("rec_location") << '\t' << r.get("rec_component") << '\t' << p.hdr.npts
             << '\t' << fmt_stream(p.hdr.delta) << '\t' << r.get("dist") << '\t' << r.get("az")
             << '\t' << r.get("baz") << '\t' << pack_path_ << '\t' << off << '\t' << bytes << '\t'
             << (kTf ? "tfpws_pack" : "pws_pack") << '\t' << final_pair_path(r) << '\n';
      ++written_;
      progress_->add("overall", 1, "worker=0 gpu=" + a_.gpu.substr(0, a_.gpu.find(',')) + " " +
                                       r.path_id + ":" + r.get("src_component") + "-" +
                                       r.get("rec_component"));
    }
    batch_.clear();
    groups_ = 0;
    return (bool)shard_;
  }
  bool close() {
    bool ok = flush();
    if (pack_) ok = fclose(pack_) == 0 && ok;
    pack_ = nullptr;
    shard_.close();
    return ok;
  }
  size_t written() const { return written_; }
  Progress *progress_ = nullptr;
  std::string pack_path_, shard_path_;

 private:
  bool fits(const Pair &p) const {
    // device bytes of the batch with p added: rows + weights + linear + out (+ spectra)
    size_t g = groups_ + p.ngroups, np = batch_.size() + 1;
    const size_t bytes = (g * p.n * 4 + g * 4 + 2 * np * p.n * 4) * 4;
    return bytes <= budget_;
  }
  bool ensure(float **p, size_t *cap, size_t n) {
    if (*cap >= n) return true;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    HIPX(hipMalloc((void **)p, std::max<size_t>(n, 1) * 4));
    *cap = n;
    return true;
  }
  const Args &a_;
  std::string out_;
  hipStream_t st_ = nullptr;
  size_t budget_ = 0, groups_ = 0, written_ = 0;
  FILE *pack_ = nullptr;
  std::ofstream shard_;
  std::vector<Pair> batch_;
  float *d_rows_ = nullptr, *d_w_ = nullptr, *d_lin_ = nullptr, *d_out_ = nullptr;
  size_t cap_rows_ = 0, cap_w_ = 0, cap_lin_ = 0, cap_out_ = 0;
};

int run(const Args &a) {
  std::vector<std::string> indexes;
  if (!read_list(a.list, &indexes)) return 1;
  const std::string out = absolute_path(a.out);
  if (mkdir_p(out) != 0) {
    FXH_LOG_ERROR("output_dir_create_failed", "path=\"%s\"", out.c_str());
    return 1;
  }
  size_t total = 0;
  {
    GroupReader counter(indexes);
    if (!counter.open()) return 1;
    std::vector<Record> g;
    while (counter.next_group(&g)) ++total;
  }
  Progress prog;
  prog.init(a.progress.empty() ? nullptr : a.progress.c_str());
  prog.set_rows({{"overall", "RUNNING", 0, total, "groups", ""}});
  Runner runner(a, out);
  runner.progress_ = &prog;
  GroupReader reader(indexes);
  if (!runner.init() || !reader.open()) {
    prog.finish("FAILED", false);
    return 1;
  }
  std::vector<Record> g;
  while (reader.next_group(&g)) {
    Pair p;
    if (!prepare_pair(g, a.substack, &p) || !runner.add(std::move(p))) {
      prog.finish("FAILED", false);
      return 1;
    }
  }
  if (!runner.close()) {
    prog.finish("FAILED", false);
    return 1;
  }
  // merged index: rows sorted by (path_id text, slot), stable within
  std::ifstream in(runner.shard_path_);
  std::string line;
  std::getline(in, line);
  struct Row {
    std::string pid;
    int slot;
    size_t seq;
    std::string line;
  };
  std::vector<Row> rows;
  while (std::getline(in, line)) {
    line = trim(line);
    if (line.empty()) continue;
    const auto f = split_tab(line);
    rows.push_back({trim(f[1]), atoi(trim(f[2]).c_str()), rows.size(), line});
  }
  std::stable_sort(rows.begin(), rows.end(), [](const Row &x, const Row &y) {
    if (x.pid != y.pid) return x.pid < y.pid;
    return x.slot < y.slot;
  });
  const std::string index_path = path_join(out, "sourcepack_index.tsv");
  std::ofstream idx(index_path);
  idx << kIndexHeader << '\n';
  for (const auto &r : rows) idx << r.line << '\n';
  idx.close();
  std::ofstream ok(path_join(out, "_SUCCESS"));
  ok << "records\t" << runner.written() << "\n"
     << "workers\t" << 1 << "\n"
     << "index\t" << index_path << "\n"
     << "pack\t" << runner.pack_path_ << "\n";
  ok.close();
  if (!idx || !ok) {
    prog.finish("FAILED", false);
    return 1;
  }
  FXH_LOG_INFO(kTf ? "tfpws_completed" : "pws_sourcepack_completed", "completed=%zu output=\"%s\"",
               runner.written(), out.c_str());
  prog.finish("DONE", true);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  Args a;
  if (!parse_args(argc, argv, &a)) return 1;
  return run(a);
}
