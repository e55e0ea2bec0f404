// SourcePack indexer over xcpack output (fastxc/operators/sourcepack/builder.py).
#include "sourcepack.hpp"

#include "fxc_host.hpp"

#include <dirent.h>
#include <ftw.h>
#include <limits.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <deque>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <string_view>
#include <tuple>

namespace fxh {

namespace {

// SOURCEPACK_INDEX_HEADER (fastxc/io/sourcepack.py:30-56)
const char *const kHeader[] = {
    "timestamp", "path_id", "component_slot", "source_key", "receiver_key", "src_id", "rec_id",
    "src_network", "src_station", "src_location", "src_component", "rec_network", "rec_station",
    "rec_location", "rec_component", "npts", "dt", "dist", "az", "baz", "record_path",
    "record_offset", "bytes", "storage_kind", "final_pair_path"};
constexpr size_t kNumColumns = sizeof(kHeader) / sizeof(kHeader[0]);

// XC_PACK_REQUIRED_FIELDS (fastxc/io/sourcepack.py:8-28)
const char *const kRequired[] = {"timestamp", "pack_path", "offset", "bytes", "src_id",
                                 "rec_id", "src_network", "src_station", "src_location",
                                 "src_component", "rec_network", "rec_station", "rec_location",
                                 "rec_component", "npts", "dt", "dist", "az", "baz"};

// the xcpack columns the index reads, in this order
enum Field {
  F_TIMESTAMP, F_PATH_ID, F_SRC_ID, F_REC_ID, F_SRC_NET, F_SRC_STA, F_SRC_LOC, F_SRC_CMP,
  F_REC_NET, F_REC_STA, F_REC_LOC, F_REC_CMP, F_NPTS, F_DT, F_DIST, F_AZ, F_BAZ, F_PACK,
  F_OFFSET, F_BYTES, F_FINAL, F_COUNT
};
const char *const kFieldNames[F_COUNT] = {
    "timestamp", "path_id", "src_id", "rec_id", "src_network", "src_station", "src_location",
    "src_component", "rec_network", "rec_station", "rec_location", "rec_component", "npts", "dt",
    "dist", "az", "baz", "pack_path", "offset", "bytes", "final_pair_path"};

// One xcpack TSV read whole; the records' fields point into it.
struct TsvData {
  std::string text;
  std::deque<std::string> unquoted;             // quoted fields, quotes removed
  std::map<std::string, std::string> paths;     // pack_path text -> resolved path
};

struct XcRecord {
  long long sequence = 0;
  std::string_view v[F_COUNT];  // absent columns and short rows read as ""
  std::string_view pack_path;
  long long offset = 0, bytes = 0;
  // sort and row values, filled once
  std::string path_id, src_key, rec_key;
  int slot = 0;

  std::string_view get(Field f) const { return v[f]; }
};

std::string_view trim_view(std::string_view s) {
  size_t a = 0, b = s.size();
  while (a < b && isspace((unsigned char)s[a])) ++a;
  while (b > a && isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

// "E"/"1" 0, "N"/"2" 1, "Z"/"3" 2, "R" 3, "T" 4, else 100, on the last
// character upper-cased (builder.py:412-418)
int component_rank(std::string_view c) {
  if (c.empty()) return 100;
  switch (toupper((unsigned char)c.back())) {
    case 'E': case '1': return 0;
    case 'N': case '2': return 1;
    case 'Z': case '3': return 2;
    case 'R': return 3;
    case 'T': return 4;
    default: return 100;
  }
}

// path_id column zero-filled to 8 when all digits, else src_id*10000+rec_id
// as %08d (builder.py:425-429)
std::string path_id_text(const XcRecord &r) {
  const std::string_view pid = trim_view(r.get(F_PATH_ID));
  if (!pid.empty()) {
    const bool digits =
        std::all_of(pid.begin(), pid.end(), [](char c) { return isdigit((unsigned char)c); });
    if (digits && pid.size() < 8) return std::string(8 - pid.size(), '0') + std::string(pid);
    return std::string(pid);
  }
  char buf[48];
  snprintf(buf, sizeof(buf), "%08lld",
           atoll(std::string(r.get(F_SRC_ID)).c_str()) * 10000LL +
               atoll(std::string(r.get(F_REC_ID)).c_str()));
  return buf;
}

// One field as Python's csv writer writes it with a tab delimiter and minimal
// quoting: quoted (quotes doubled) when it holds a tab, quote, CR or LF.
void append_csv_field(std::string *out, std::string_view v) {
  if (v.find_first_of("\t\"\r\n") == std::string_view::npos) {
    out->append(v.data(), v.size());
    return;
  }
  *out += '"';
  for (char c : v) {
    if (c == '"') *out += '"';
    *out += c;
  }
  *out += '"';
}

// One line split as Python's csv reader does with a tab delimiter: a field
// that starts with a quote runs to the closing quote ("" is a quote).
void csv_split(std::string_view line, std::vector<std::string_view> *out,
               std::deque<std::string> *unquoted) {
  out->clear();
  size_t i = 0;
  for (;;) {
    if (i < line.size() && line[i] == '"') {
      std::string f;
      ++i;
      while (i < line.size()) {
        if (line[i] == '"') {
          if (i + 1 < line.size() && line[i + 1] == '"') {
            f += '"';
            i += 2;
            continue;
          }
          ++i;
          break;
        }
        f += line[i++];
      }
      while (i < line.size() && line[i] != '\t') f += line[i++];
      unquoted->push_back(std::move(f));
      out->push_back(unquoted->back());
    } else {
      size_t j = line.find('\t', i);
      if (j == std::string_view::npos) j = line.size();
      out->push_back(line.substr(i, j - i));
      i = j;
    }
    if (i >= line.size()) break;
    ++i;  // the tab
  }
}

std::vector<std::string> csv_split_owned(const std::string &line) {
  std::vector<std::string_view> v;
  std::deque<std::string> uq;
  csv_split(line, &v, &uq);
  return std::vector<std::string>(v.begin(), v.end());
}

// pathlib's spelling of a path: "." parts and repeated or trailing slashes
// dropped (".." kept), "~/" expanded.
std::string pathlib_text(const std::string &p) {
  std::string s = p;
  if (s == "~" || s.compare(0, 2, "~/") == 0) {
    const char *home = getenv("HOME");
    if (home) s = std::string(home) + s.substr(1);
  }
  const bool abs = !s.empty() && s[0] == '/';
  std::string out;
  size_t i = 0;
  while (i <= s.size()) {
    size_t j = s.find('/', i);
    if (j == std::string::npos) j = s.size();
    const std::string part = s.substr(i, j - i);
    if (!part.empty() && part != ".") out += (out.empty() && !abs ? "" : "/") + part;
    i = j + 1;
  }
  if (out.empty()) return abs ? "/" : ".";
  return out;
}

std::string resolved(const std::string &p) {
  char buf[PATH_MAX];
  if (realpath(pathlib_text(p).c_str(), buf)) return buf;
  return pathlib_text(absolute_path(pathlib_text(p)));
}

bool parse_int(std::string_view text, long long *out) {
  const std::string t(trim_view(text));
  if (t.empty()) return false;
  char *end = nullptr;
  *out = strtoll(t.c_str(), &end, 10);
  return *end == '\0';
}

// read_xcpack_tsv (fastxc/io/sourcepack.py:120-154): header check, relative
// pack paths against the TSV's folder, each must exist, offset >= 0, bytes > 0.
bool read_xcpack_tsv(const std::string &tsv, long long first_seq, TsvData *data,
                     std::vector<XcRecord> *out) {
  {
    std::ifstream in(tsv, std::ios::binary);
    if (!in) {
      FXH_LOG_ERROR("xcpack_tsv_read_failed", "path=\"%s\"", tsv.c_str());
      return false;
    }
    in.seekg(0, std::ios::end);
    data->text.resize((size_t)in.tellg());
    in.seekg(0);
    in.read(&data->text[0], (std::streamsize)data->text.size());
  }
  std::string_view all(data->text);
  if (all.substr(0, 3) == "\xEF\xBB\xBF") all.remove_prefix(3);
  auto next_line = [&all](std::string_view *line) {
    if (all.empty()) return false;
    size_t e = all.find('\n');
    if (e == std::string_view::npos) e = all.size();
    *line = all.substr(0, e);
    all.remove_prefix(e < all.size() ? e + 1 : e);
    if (!line->empty() && line->back() == '\r') line->remove_suffix(1);
    return true;
  };
  std::string_view line;
  if (!next_line(&line)) return true;  // empty file: no records
  std::vector<std::string_view> fields;
  csv_split(line, &fields, &data->unquoted);
  const std::vector<std::string> names(fields.begin(), fields.end());
  for (const char *r : kRequired)
    if (column_index(names, r) < 0) {
      FXH_LOG_ERROR("xcpack_tsv_missing_field", "path=\"%s\" field=%s", tsv.c_str(), r);
      return false;
    }
  int col[F_COUNT];
  for (int f = 0; f < F_COUNT; ++f) col[f] = column_index(names, kFieldNames[f]);
  const std::string dir = dirname_of(tsv);
  long long seq = first_seq;
  while (next_line(&line)) {
    if (line.empty()) continue;
    XcRecord r;
    csv_split(line, &fields, &data->unquoted);
    for (int f = 0; f < F_COUNT; ++f)
      if (col[f] >= 0 && (size_t)col[f] < fields.size()) r.v[f] = fields[col[f]];
    r.sequence = seq;
    const std::string text(r.get(F_PACK));
    auto hit = data->paths.find(text);
    if (hit == data->paths.end()) {
      std::string p = pathlib_text(text);
      if (p[0] != '/') p = pathlib_text(path_join(dir, p));
      if (!is_file(p)) {
        FXH_LOG_ERROR("xcpack_file_missing", "tsv=\"%s\" pack=\"%s\"", tsv.c_str(), p.c_str());
        return false;
      }
      hit = data->paths.emplace(text, p).first;
    }
    r.pack_path = hit->second;
    if (!parse_int(r.get(F_OFFSET), &r.offset) || !parse_int(r.get(F_BYTES), &r.bytes) ||
        r.offset < 0 || r.bytes <= 0) {
      FXH_LOG_ERROR("xcpack_offset_invalid", "tsv=\"%s\" offset=\"%s\" bytes=\"%s\"",
                    tsv.c_str(), std::string(r.get(F_OFFSET)).c_str(),
                    std::string(r.get(F_BYTES)).c_str());
      return false;
    }
    r.path_id = path_id_text(r);
    r.slot = component_rank(r.get(F_SRC_CMP)) * 3 + component_rank(r.get(F_REC_CMP));
    for (int side = 0; side < 2; ++side) {
      std::string &k = side ? r.rec_key : r.src_key;
      const Field net = side ? F_REC_NET : F_SRC_NET, sta = side ? F_REC_STA : F_SRC_STA,
                  loc = side ? F_REC_LOC : F_SRC_LOC;
      k.reserve(r.v[net].size() + r.v[sta].size() + r.v[loc].size() + 2);
      k.append(r.v[net]).append(1, '.').append(r.v[sta]).append(1, '.').append(r.v[loc]);
    }
    ++seq;
    out->push_back(std::move(r));
  }
  return true;
}

std::vector<std::string> sorted_tsvs(const std::string &xcpack_dir) {
  auto v = list_files_with_suffix(xcpack_dir, ".tsv");
  std::sort(v.begin(), v.end());
  return v;
}

int rm_entry(const char *path, const struct stat *, int, struct FTW *) { return ::remove(path); }

void rmtree(const std::string &p) {
  struct stat st;
  if (lstat(p.c_str(), &st) == 0) nftw(p.c_str(), rm_entry, 16, FTW_DEPTH | FTW_PHYS);
}

// _build_one_timestamp (builder.py:267-306): <out>/.<safe>.tmp holds the
// index and _SUCCESS, then replaces <out>/<safe>.
bool build_one(const std::string &out_dir, const std::string &ts, std::vector<XcRecord *> recs,
               bool sort_within_source, SourcePackResult *res) {
  const std::string safe = sourcepack_safe_name(ts);
  const std::string target = path_join(out_dir, safe);
  const std::string tmp = path_join(out_dir, "." + safe + ".tmp");
  rmtree(tmp);
  if (mkdir_p(tmp) != 0) {
    FXH_LOG_ERROR("sourcepack_dir_create_failed", "path=\"%s\"", tmp.c_str());
    return false;
  }
  if (sort_within_source) {
    std::sort(recs.begin(), recs.end(), [](const XcRecord *a, const XcRecord *b) {
      return std::tie(a->path_id, a->slot, a->src_key, a->rec_key, a->sequence) <
             std::tie(b->path_id, b->slot, b->src_key, b->rec_key, b->sequence);
    });
  } else {
    std::sort(recs.begin(), recs.end(),
              [](const XcRecord *a, const XcRecord *b) { return a->sequence < b->sequence; });
  }
  std::set<std::string> sources;
  uint64_t nbytes = 0;
  std::string text;
  for (size_t i = 0; i < kNumColumns; ++i) (text += i ? "\t" : "") += kHeader[i];
  text += "\r\n";
  for (const XcRecord *r : recs) {
    sources.insert(r->src_key);
    nbytes += (uint64_t)r->bytes;
    const std::string_view vals[kNumColumns] = {
        r->v[F_TIMESTAMP], r->path_id, std::string_view(), r->src_key, r->rec_key, r->v[F_SRC_ID],
        r->v[F_REC_ID], r->v[F_SRC_NET], r->v[F_SRC_STA], r->v[F_SRC_LOC], r->v[F_SRC_CMP],
        r->v[F_REC_NET], r->v[F_REC_STA], r->v[F_REC_LOC], r->v[F_REC_CMP], r->v[F_NPTS],
        r->v[F_DT], r->v[F_DIST], r->v[F_AZ], r->v[F_BAZ], r->pack_path, std::string_view(),
        std::string_view(), "xcpack", r->v[F_FINAL]};
    for (size_t i = 0; i < kNumColumns; ++i) {
      if (i) text += '\t';
      if (i == 2) text += std::to_string(r->slot);
      else if (i == 21) text += std::to_string(r->offset);
      else if (i == 22) text += std::to_string(r->bytes);
      else append_csv_field(&text, vals[i]);
    }
    text += "\r\n";
  }
  bool ok;
  {
    std::ofstream f(path_join(tmp, "sourcepack_index.tsv"), std::ios::binary);
    f << text;
    std::ofstream s(path_join(tmp, "_SUCCESS"), std::ios::binary);
    s << "timestamp\t" << ts << "\nrecords\t" << recs.size() << "\nsources\t" << sources.size()
      << "\nbytes_indexed\t" << nbytes << "\n";
    f.close();
    s.close();
    ok = f.good() && s.good();
  }
  if (!ok) {
    FXH_LOG_ERROR("sourcepack_write_failed", "path=\"%s\"", tmp.c_str());
    return false;
  }
  rmtree(target);
  if (::rename(tmp.c_str(), target.c_str()) != 0) {
    FXH_LOG_ERROR("sourcepack_rename_failed", "from=\"%s\" to=\"%s\"", tmp.c_str(),
                  target.c_str());
    return false;
  }
  res->index_paths.push_back(path_join(target, "sourcepack_index.tsv"));
  res->records += (long long)recs.size();
  res->sources += (long long)sources.size();
  res->bytes_indexed += nbytes;
  return true;
}

// manifest/timestamp_index.tsv of the workspace holding xcpack (directly or
// under ncf/), its non-empty "timestamp" values (builder.py:362-382)
void manifest_timestamps(const std::string &xcpack_dir, std::set<std::string> *out) {
  const std::string parent = dirname_of(xcpack_dir);
  const std::string pname = parent.substr(parent.rfind('/') + 1);
  const std::string ws = pname == "ncf" ? dirname_of(parent) : parent;
  std::ifstream man(path_join(ws, "manifest/timestamp_index.tsv"), std::ios::binary);
  std::string line;
  if (!man || !std::getline(man, line)) return;
  if (line.size() >= 3 && line.compare(0, 3, "\xEF\xBB\xBF") == 0) line = line.substr(3);
  if (!line.empty() && line.back() == '\r') line.pop_back();
  const int col = column_index(csv_split_owned(line), "timestamp");
  if (col < 0) return;
  while (std::getline(man, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    const auto f = csv_split_owned(line);
    if ((int)f.size() > col && !trim(f[col]).empty()) out->insert(trim(f[col]));
  }
}

}  // namespace

std::string xcpack_dir_of(const std::string &root) {
  const std::string r = resolved(root);
  const std::string name = r.substr(r.rfind('/') + 1);
  return name == "xcpack" ? r : path_join(r, "xcpack");
}

std::string sourcepack_safe_name(const std::string &text) {
  std::string o;
  bool run = false;
  for (char c : text) {
    const bool ok = isalnum((unsigned char)c) || c == '.' || c == '_' || c == ':' || c == '-';
    if (ok) {
      o += c;
      run = false;
    } else if (!run) {
      o += '-';
      run = true;
    }
  }
  const size_t a = o.find_first_not_of('-');
  if (a == std::string::npos) return "item";
  const size_t b = o.find_last_not_of('-');
  return o.substr(a, b - a + 1);
}

// build_sourcepack_timestamp (builder.py:242-264): TSVs whose names start with
// the timestamp (else all), sequence counted over every row read.
bool sourcepack_build_timestamp(const std::string &xcpack_dir, const std::string &out_dir,
                                const std::string &ts, bool sort_within_source,
                                SourcePackResult *result) {
  if (!is_dir(xcpack_dir)) {
    FXH_LOG_ERROR("xcpack_dir_missing", "path=\"%s\"", xcpack_dir.c_str());
    return false;
  }
  auto all = sorted_tsvs(xcpack_dir);
  std::vector<std::string> tsvs;
  for (const auto &p : all) {
    const std::string name = p.substr(p.rfind('/') + 1);
    if (name.compare(0, ts.size(), ts) == 0) tsvs.push_back(p);
  }
  if (tsvs.empty()) tsvs = all;
  std::deque<TsvData> data;
  std::vector<XcRecord> recs;
  for (const auto &t : tsvs) {
    data.emplace_back();
    if (!read_xcpack_tsv(t, (long long)recs.size(), &data.back(), &recs)) return false;
  }
  std::vector<XcRecord *> mine;
  for (auto &r : recs)
    if (r.get(F_TIMESTAMP) == ts) mine.push_back(&r);
  if (mkdir_p(out_dir) != 0) {
    FXH_LOG_ERROR("sourcepack_dir_create_failed", "path=\"%s\"", out_dir.c_str());
    return false;
  }
  return build_one(out_dir, ts, std::move(mine), sort_within_source, result);
}

// build_sourcepack (builder.py:163-239) without its progress calls.
bool sourcepack_build_all(const std::string &xcpack_dir, const std::string &out_dir,
                          bool sort_within_source, SourcePackResult *result) {
  if (!is_dir(xcpack_dir)) {
    FXH_LOG_ERROR("xcpack_dir_missing", "path=\"%s\"", xcpack_dir.c_str());
    return false;
  }
  if (mkdir_p(out_dir) != 0) {
    FXH_LOG_ERROR("sourcepack_dir_create_failed", "path=\"%s\"", out_dir.c_str());
    return false;
  }
  const auto tsvs = sorted_tsvs(xcpack_dir);
  std::deque<TsvData> data;
  std::vector<XcRecord> recs;
  for (const auto &t : tsvs) {
    data.emplace_back();
    if (!read_xcpack_tsv(t, (long long)recs.size(), &data.back(), &recs)) return false;
  }
  std::map<std::string, std::vector<XcRecord *>> by_ts;
  for (auto &r : recs) by_ts[std::string(r.get(F_TIMESTAMP))].push_back(&r);
  std::set<std::string> stamps;
  for (const auto &kv : by_ts) stamps.insert(kv.first);
  manifest_timestamps(xcpack_dir, &stamps);
  for (const auto &p : list_files_with_suffix(xcpack_dir, ".done")) {
    const std::string name = p.substr(p.rfind('/') + 1);
    if (name[0] != '_') stamps.insert(name.substr(0, name.size() - 5));
  }
  if (stamps.empty()) {
    FXH_LOG_ERROR("xcpack_timestamps_missing", "path=\"%s\"", xcpack_dir.c_str());
    return false;
  }
  for (const auto &ts : stamps) {
    auto it = by_ts.find(ts);
    std::vector<XcRecord *> group;
    if (it != by_ts.end()) group = std::move(it->second);
    if (!build_one(out_dir, ts, std::move(group), sort_within_source, result)) return false;
  }
  return true;
}

// AsyncSourcePackMaterializer.finish (builder.py:87-95): the count of
// <out>/*/sourcepack_index.tsv and the summed per-timestamp figures.
bool sourcepack_write_success(const std::string &out_dir, const SourcePackResult &total) {
  size_t n = 0;
  if (DIR *d = opendir(out_dir.c_str())) {
    while (struct dirent *e = readdir(d)) {
      const std::string name = e->d_name;
      if (name == "." || name == "..") continue;
      if (is_file(path_join(path_join(out_dir, name), "sourcepack_index.tsv"))) ++n;
    }
    closedir(d);
  }
  const std::string tmp = path_join(out_dir, "._SUCCESS.tmp");
  {
    std::ofstream s(tmp, std::ios::binary);
    s << "timestamps\t" << n << "\nrecords\t" << total.records << "\nsources\t" << total.sources
      << "\nbytes_indexed\t" << total.bytes_indexed << "\n";
    s.close();
    if (!s.good()) return false;
  }
  return ::rename(tmp.c_str(), path_join(out_dir, "_SUCCESS").c_str()) == 0;
}

size_t sourcepack_expected_count(const std::string &xcpack_dir) {
  std::set<std::string> stamps;
  manifest_timestamps(xcpack_dir, &stamps);
  if (!stamps.empty()) return stamps.size();
  size_t n = 0;
  for (const auto &p : list_files_with_suffix(xcpack_dir, ".done"))
    if (p[p.rfind('/') + 1] != '_') ++n;
  return n;
}

SourcePackMaterializer::SourcePackMaterializer(const std::string &xc_output_root,
                                               const std::string &sourcepack_dir,
                                               bool sort_within_source,
                                               const std::string &progress_file)
    : xcpack_dir_(xcpack_dir_of(xc_output_root)),
      out_dir_(resolved(sourcepack_dir)),
      sort_(sort_within_source) {
  if (!progress_file.empty() && progress_file != "NONE") mkdir_p(dirname_of(progress_file));
  progress_.init(progress_file.c_str());
  expected_ = sourcepack_expected_count(xcpack_dir_);
}

SourcePackMaterializer::~SourcePackMaterializer() {
  // the producer failed before finish: what was handed over is built, the
  // global marker is withheld
  if (started_ && !finished_) finish(false, "upstream XC did not complete; global success marker withheld");
}

void SourcePackMaterializer::start() {
  started_ = true;
  mkdir_p(out_dir_);
  progress_.update("overall", "RUNNING", 0, expected_, "timestamps", "async source index build");
  thread_ = std::thread([this] { run(); });
}

void SourcePackMaterializer::submit(const std::string &timestamp) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    queue_.push_back(timestamp);
  }
  cv_.notify_all();
}

bool SourcePackMaterializer::build(const std::string &ts) {
  SourcePackResult one;
  if (!sourcepack_build_timestamp(xcpack_dir_, out_dir_, ts, sort_, &one)) {
    std::lock_guard<std::mutex> lk(mu_);
    failed_ = true;
    if (first_error_.empty()) first_error_ = "source index build failed for " + ts;
    return false;
  }
  size_t completed;
  {
    std::lock_guard<std::mutex> lk(mu_);
    done_.insert(ts);
    total_.records += one.records;
    total_.sources += one.sources;
    total_.bytes_indexed += one.bytes_indexed;
    total_.index_paths.insert(total_.index_paths.end(), one.index_paths.begin(),
                              one.index_paths.end());
    completed = done_.size();
  }
  progress_.update("overall", "RUNNING", completed, expected_, "timestamps",
                   ts + ": " + std::to_string(one.records) + " records");
  FXH_LOG_INFO("sourcepack_timestamp_done", "timestamp=\"%s\" records=%lld sources=%lld",
               ts.c_str(), one.records, one.sources);
  return true;
}

void SourcePackMaterializer::run() {
  for (;;) {
    std::string ts;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [this] { return stop_ || !queue_.empty(); });
      if (queue_.empty()) return;  // stop with nothing left
      ts = queue_.front();
      queue_.pop_front();
      if (done_.count(ts)) continue;
    }
    if (!build(ts)) return;  // the first failure ends the thread; finish reports it
  }
}

bool SourcePackMaterializer::finish(bool mark_success, const std::string &status_message) {
  finished_ = true;
  if (thread_.joinable()) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    thread_.join();
  }
  // final scan: every <ts>.done not built yet, in name order
  if (!failed_ && is_dir(xcpack_dir_)) {
    auto marks = list_files_with_suffix(xcpack_dir_, ".done");
    std::sort(marks.begin(), marks.end());
    for (const auto &p : marks) {
      const std::string name = p.substr(p.rfind('/') + 1);
      if (name[0] == '_') continue;
      const std::string ts = name.substr(0, name.size() - 5);
      if (done_.count(ts)) continue;
      if (!build(ts)) break;
    }
  }
  if (failed_) {
    progress_.update("overall", "FAILED", done_.size(), expected_, "timestamps", first_error_);
    FXH_LOG_ERROR("sourcepack_failed", "detail=\"%s\"", first_error_.c_str());
    return false;
  }
  size_t nidx = 0;
  if (DIR *d = opendir(out_dir_.c_str())) {
    while (struct dirent *e = readdir(d)) {
      const std::string name = e->d_name;
      if (name != "." && name != ".." &&
          is_file(path_join(path_join(out_dir_, name), "sourcepack_index.tsv")))
        ++nidx;
    }
    closedir(d);
  }
  const size_t total = expected_ ? expected_ : nidx;
  if (mark_success) {
    if (!sourcepack_write_success(out_dir_, total_)) {
      FXH_LOG_ERROR("sourcepack_success_write_failed", "path=\"%s\"", out_dir_.c_str());
      return false;
    }
    progress_.update("overall", "DONE", nidx, total, "timestamps", "source index ready");
  } else {
    ::remove(path_join(out_dir_, "_SUCCESS").c_str());
    progress_.update("overall", "FAILED", nidx, total, "timestamps",
                     status_message.empty()
                         ? "upstream XC did not complete; global success marker withheld"
                         : status_message);
  }
  return true;
}

}  // namespace fxh
