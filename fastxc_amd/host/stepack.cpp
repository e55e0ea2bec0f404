#include "stepack.hpp"

#include "fxc_host.hpp"

#include <errno.h>
#include <fcntl.h>
#include <inttypes.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>

namespace fxh {

namespace {

const char kTsvHeader[] =
    "timestamp\tworker_id\tbatch_seq\tstart_group\tgroup_count\t"
    "pack_path\toffset\tbytes\tversion\tlayout\tnstep\tnslc_start\t"
    "nslc_count\tbatch_nslc_count\tnspec\tdt\tdf\tnslc_table_bytes\t"
    "payload_offset\tpayload_bytes\tbatch_payload_bytes\tstep_bytes\t"
    "pitch_step_bytes\tnslc_step_bytes\tfirst_nsl_id\tlast_nsl_id\n";

// NUL-padded copy with tabs/newlines blanked (fixed-width header strings).
void fixed_copy(char *dst, size_t size, const std::string &s) {
  memset(dst, 0, size);
  for (size_t i = 0; i + 1 < size && i < s.size(); ++i) {
    const char c = s[i];
    dst[i] = (c == '\t' || c == '\n' || c == '\r') ? ' ' : c;
  }
}

std::string fixed_read(const char *src, size_t size) {
  size_t n = 0;
  while (n < size && src[n]) ++n;
  return std::string(src, n);
}

}  // namespace

int stepack_write_batch(const std::string &root, int worker_id, size_t batch_seq,
                        size_t start_group, size_t group_count,
                        const std::vector<StepackNode> &nodes, int nstep, int nspec, float dt,
                        float df, const std::complex<float> *spectra) {
  if (nodes.empty()) return 0;
  if (mkdir_p(root) != 0) {
    FXH_LOG_ERROR("stepack_dir_create_failed", "path=\"%s\"", root.c_str());
    return -1;
  }
  const uint64_t n = nodes.size();
  const uint64_t table_bytes = n * sizeof(StepackNslc);
  const uint64_t nslc_step_bytes = (uint64_t)nspec * 8;
  const uint64_t pitch_step_bytes = n * nslc_step_bytes;
  const uint64_t payload_bytes = (uint64_t)nstep * pitch_step_bytes;
  const uint64_t payload_offset = sizeof(StepackHeader) + table_bytes;
  const uint64_t record_bytes = payload_offset + payload_bytes;
  size_t num_ch = 0;
  if (group_count > 0 && n % group_count == 0) num_ch = n / group_count;

  StepackHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, "FXCSTPK", 7);
  h.version = kStepackVersion;
  h.header_size = sizeof(StepackHeader);
  h.nslc_entry_size = sizeof(StepackNslc);
  h.layout = kStepackLayout;
  h.batch_seq = batch_seq;
  h.start_group = start_group;
  h.group_count = group_count;
  h.worker_id = (uint32_t)worker_id;
  h.nstep = (uint32_t)nstep;
  h.nslc_count = (uint32_t)n;
  h.nspec = (uint32_t)nspec;
  h.dt = dt;
  h.df = df;
  h.nslc_table_bytes = table_bytes;
  h.payload_offset = payload_offset;
  h.payload_bytes = payload_bytes;
  h.pitch_step_bytes = pitch_step_bytes;
  fixed_copy(h.first_timestamp, sizeof(h.first_timestamp), nodes[0].timestamp);

  std::vector<StepackNslc> table(n);
  for (size_t i = 0; i < n; ++i) {
    StepackNslc &e = table[i];
    memset(&e, 0, sizeof(e));
    e.file_index = (uint32_t)i;
    e.nsl_id = nodes[i].gnsl_id;
    e.stla = nodes[i].stla;
    e.stlo = nodes[i].stlo;
    fixed_copy(e.network, sizeof(e.network), nodes[i].network);
    fixed_copy(e.station, sizeof(e.station), nodes[i].station);
    fixed_copy(e.location, sizeof(e.location), nodes[i].location);
    fixed_copy(e.component, sizeof(e.component), nodes[i].component);
  }

  char leaf[64];
  snprintf(leaf, sizeof(leaf), "w%03d.b%06zu.stepack", worker_id, batch_seq);
  const std::string pack_path = path_join(absolute_path(root), leaf);
  snprintf(leaf, sizeof(leaf), "w%03d.b%06zu.tsv", worker_id, batch_seq);
  const std::string tsv_path = path_join(absolute_path(root), leaf);

  FILE *pf = fopen(pack_path.c_str(), "wb");
  if (!pf) {
    FXH_LOG_ERROR("stepack_open_failed", "path=\"%s\" error=\"%s\"", pack_path.c_str(),
                  strerror(errno));
    return -1;
  }
  bool ok = fwrite(&h, sizeof(h), 1, pf) == 1 &&
            fwrite(table.data(), table_bytes, 1, pf) == 1 &&
            (payload_bytes == 0 || fwrite(spectra, payload_bytes, 1, pf) == 1);
  ok = (fclose(pf) == 0) && ok;
  if (!ok) {
    FXH_LOG_ERROR("stepack_batch_write_failed", "path=\"%s\" bytes=%" PRIu64, pack_path.c_str(),
                  record_bytes);
    return -1;
  }

  FILE *tf = fopen(tsv_path.c_str(), "w");
  if (!tf) {
    FXH_LOG_ERROR("stepack_tsv_open_failed", "path=\"%s\"", tsv_path.c_str());
    return -1;
  }
  fputs(kTsvHeader, tf);
  // one row per run of nodes sharing a timestamp
  size_t begin = 0;
  while (begin < n) {
    size_t end = begin + 1;
    while (end < n && nodes[end].timestamp == nodes[begin].timestamp) ++end;
    const size_t run = end - begin;
    size_t run_start_group = start_group, run_group_count = group_count;
    if (num_ch > 0 && begin % num_ch == 0 && run % num_ch == 0) {
      run_start_group = start_group + begin / num_ch;
      run_group_count = run / num_ch;
    }
    const uint64_t run_step_bytes = run * nslc_step_bytes;
    const uint64_t run_payload_bytes = (uint64_t)nstep * run_step_bytes;
    fprintf(tf, "%s\t%d\t%zu\t%zu\t%zu\t%s", clean_tsv_field(nodes[begin].timestamp).c_str(),
            worker_id, batch_seq, run_start_group, run_group_count,
            clean_tsv_field(pack_path).c_str());
    fprintf(tf,
            "\t%" PRIu64 "\t%" PRIu64 "\t%u\t%u\t%d\t%zu\t%zu\t%zu\t%d\t%.9g\t%.9g\t%" PRIu64
            "\t%" PRIu64 "\t%" PRIu64 "\t%" PRIu64 "\t%" PRIu64 "\t%" PRIu64 "\t%" PRIu64 "\t",
            (uint64_t)0, record_bytes, kStepackVersion, kStepackLayout, nstep, begin, run,
            (size_t)n, nspec, (double)dt, (double)df, table_bytes, payload_offset,
            run_payload_bytes, payload_bytes, run_step_bytes, pitch_step_bytes, nslc_step_bytes);
    fprintf(tf, "%s\t%s\n", clean_tsv_field(nodes[begin].nsl_id).c_str(),
            clean_tsv_field(nodes[end - 1].nsl_id).c_str());
    begin = end;
  }
  if (fclose(tf) != 0) {
    FXH_LOG_ERROR("close_stepack_tsv_failed", "path=\"%s\"", tsv_path.c_str());
    return -1;
  }
  return 0;
}

// ----------------------------------------------------------------- reading --
namespace {

bool parse_fragment_row(const std::vector<std::string> &f, const std::vector<std::string> &hdr,
                        const std::string &base_dir, StepackFragment *out, std::string *ts) {
  auto col = [&](const char *name) { return column_index(hdr, name); };
  const int c_ts = col("timestamp"), c_path = col("pack_path"), c_nstep = col("nstep"),
            c_start = col("nslc_start"), c_count = col("nslc_count"),
            c_batch = col("batch_nslc_count"), c_nspec = col("nspec"), c_dt = col("dt"),
            c_df = col("df"), c_poff = col("payload_offset"), c_pbytes = col("payload_bytes"),
            c_step = col("step_bytes"), c_pitch = col("pitch_step_bytes"),
            c_nstepb = col("nslc_step_bytes");
  const int req[] = {c_ts, c_path, c_nstep, c_start, c_count, c_batch, c_nspec,
                     c_dt, c_df, c_poff, c_pbytes, c_step, c_pitch, c_nstepb};
  for (int c : req)
    if (c < 0 || c >= (int)f.size()) return false;
  StepackFragment fr;
  *ts = normalize_timestamp_for_xc(f[c_ts]);
  if (ts->empty()) return false;
  const std::string p = trim(f[c_path]);
  fr.pack_path = (!p.empty() && p[0] == '/') ? p : path_join(base_dir, p);
  auto opt_size = [&](const char *name, size_t *dst) {
    const int c = col(name);
    return c < 0 || c >= (int)f.size() || parse_size_strict(f[c], dst);
  };
  if (!opt_size("worker_id", &fr.worker_id) || !opt_size("batch_seq", &fr.batch_seq) ||
      !opt_size("start_group", &fr.start_group) || !opt_size("group_count", &fr.group_count) ||
      !parse_size_strict(f[c_nstep], &fr.nstep) || !parse_size_strict(f[c_start], &fr.nslc_start) ||
      !parse_size_strict(f[c_count], &fr.nslc_count) ||
      !parse_size_strict(f[c_batch], &fr.batch_nslc_count) ||
      !parse_size_strict(f[c_nspec], &fr.nspec) || !parse_float_strict(f[c_dt], &fr.dt) ||
      !parse_float_strict(f[c_df], &fr.df) || !parse_size_strict(f[c_poff], &fr.payload_offset) ||
      !parse_size_strict(f[c_pbytes], &fr.payload_bytes) ||
      !parse_size_strict(f[c_step], &fr.step_bytes) ||
      !parse_size_strict(f[c_pitch], &fr.pitch_step_bytes) ||
      !parse_size_strict(f[c_nstepb], &fr.nslc_step_bytes)) {
    FXH_LOG_ERROR("stepack_tsv_row_numeric_invalid", "pack=\"%s\"", fr.pack_path.c_str());
    return false;
  }
  if (fr.nslc_count == 0 || fr.nstep == 0 || fr.nspec < 2 || fr.nslc_step_bytes != fr.nspec * 8 ||
      fr.step_bytes != fr.nslc_count * fr.nslc_step_bytes ||
      fr.payload_bytes != fr.nstep * fr.step_bytes) {
    FXH_LOG_ERROR("stepack_tsv_row_shape_invalid", "pack=\"%s\"", fr.pack_path.c_str());
    return false;
  }
  *out = fr;
  return true;
}

bool tsv_has_columns(const std::string &path, const char *a, const char *b) {
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    const std::string t = trim(line);
    if (t.empty() || t[0] == '#') continue;
    const auto h = split_tab(t);
    return column_index(h, a) >= 0 && column_index(h, b) >= 0;
  }
  return false;
}

}  // namespace

bool load_timestamp_inputs(const std::string &input_path, std::vector<TimestampInput> *out) {
  std::vector<std::string> tsvs;
  if (is_dir(input_path)) {
    std::string dir = input_path;
    if (is_dir(path_join(dir, "stepack"))) dir = path_join(dir, "stepack");
    tsvs = list_files_with_suffix(dir, ".tsv");
    if (tsvs.empty()) {
      FXH_LOG_ERROR("stepack_tsv_not_found", "path=\"%s\"", dir.c_str());
      return false;
    }
  } else if (tsv_has_columns(input_path, "pack_path", "nslc_start")) {
    tsvs.push_back(input_path);
  } else {
    FXH_LOG_ERROR("stepack_input_invalid", "path=\"%s\" expected=stepack_directory_or_stepack_tsv",
                  input_path.c_str());
    return false;
  }
  std::map<std::string, size_t> by_ts;
  out->clear();
  for (const auto &tsv : tsvs) {
    std::ifstream in(tsv);
    if (!in) {
      FXH_LOG_ERROR("stepack_tsv_open_failed", "path=\"%s\"", tsv.c_str());
      return false;
    }
    const std::string base = dirname_of(absolute_path(tsv));
    std::vector<std::string> header;
    std::string line;
    while (std::getline(in, line)) {
      const std::string t = trim(line);
      if (t.empty() || t[0] == '#') continue;
      const auto fields = split_tab(t);
      if (header.empty()) {
        header = fields;
        if (column_index(header, "pack_path") < 0 || column_index(header, "nslc_start") < 0 ||
            column_index(header, "pitch_step_bytes") < 0) {
          FXH_LOG_ERROR("stepack_tsv_columns_missing", "path=\"%s\"", tsv.c_str());
          return false;
        }
        continue;
      }
      StepackFragment fr;
      std::string ts;
      if (!parse_fragment_row(fields, header, base, &fr, &ts)) {
        FXH_LOG_ERROR("stepack_tsv_row_malformed", "path=\"%s\"", tsv.c_str());
        return false;
      }
      auto it = by_ts.find(ts);
      if (it == by_ts.end()) {
        out->push_back(TimestampInput{ts, {}});
        it = by_ts.emplace(ts, out->size() - 1).first;
      }
      (*out)[it->second].fragments.push_back(fr);
    }
  }
  std::sort(out->begin(), out->end(),
            [](const TimestampInput &a, const TimestampInput &b) { return a.timestamp < b.timestamp; });
  for (auto &ti : *out)
    std::sort(ti.fragments.begin(), ti.fragments.end(),
              [](const StepackFragment &a, const StepackFragment &b) {
                if (a.batch_seq != b.batch_seq) return a.batch_seq < b.batch_seq;
                if (a.start_group != b.start_group) return a.start_group < b.start_group;
                if (a.nslc_start != b.nslc_start) return a.nslc_start < b.nslc_start;
                return a.pack_path < b.pack_path;
              });
  FXH_LOG_INFO("stepack_index_loaded", "timestamps=%zu tsvs=%zu", out->size(), tsvs.size());
  return true;
}

bool open_timestamp_work(const TimestampInput &in, TimestampWork *w) {
  *w = TimestampWork();
  w->timestamp = in.timestamp;
  if (in.fragments.empty()) return false;
  struct Row {
    SpecMeta meta;
    size_t frag, local, order;
  };
  std::vector<Row> rows;
  for (size_t fi = 0; fi < in.fragments.size(); ++fi) {
    const StepackFragment &fr = in.fragments[fi];
    const int fd = open(fr.pack_path.c_str(), O_RDONLY);
    if (fd < 0) {
      FXH_LOG_ERROR("stepack_open_failed", "path=\"%s\" error=\"%s\"", fr.pack_path.c_str(),
                    strerror(errno));
      return false;
    }
    StepackHeader h;
    struct stat st;
    bool ok = fstat(fd, &st) == 0 && pread_exact(fd, 0, &h, sizeof(h)) == 0;
    ok = ok && memcmp(h.magic, "FXCSTPK", 7) == 0 && h.version == kStepackVersion &&
         h.header_size == sizeof(StepackHeader) && h.nslc_entry_size == sizeof(StepackNslc) &&
         h.layout == kStepackLayout && h.nstep > 0 && h.nspec >= 2 && h.nslc_count > 0 &&
         h.dt > 0.f && h.df > 0.f && h.nstep == fr.nstep && h.nspec == fr.nspec &&
         h.nslc_count == fr.batch_nslc_count && h.payload_offset == fr.payload_offset &&
         h.pitch_step_bytes == fr.pitch_step_bytes && std::fabs(h.dt - fr.dt) <= 1e-6f &&
         std::fabs(h.df - fr.df) <= 1e-9f;
    if (ok && fi > 0)
      ok = (int)h.nspec == w->nspec && (int)h.nstep == w->nstep &&
           std::fabs(h.dt - w->dt) <= 1e-6f && std::fabs(h.df - w->df) <= 1e-9f;
    ok = ok && fr.nslc_start < fr.batch_nslc_count &&
         fr.nslc_count <= fr.batch_nslc_count - fr.nslc_start &&
         h.nslc_table_bytes == (uint64_t)h.nslc_count * h.nslc_entry_size &&
         h.payload_offset >= h.header_size + h.nslc_table_bytes &&
         h.payload_offset + h.payload_bytes <= (uint64_t)st.st_size;
    if (!ok) {
      close(fd);
      FXH_LOG_ERROR("stepack_header_invalid", "path=\"%s\"", fr.pack_path.c_str());
      return false;
    }
    if (fi == 0) {
      w->nspec = (int)h.nspec;
      w->nstep = (int)h.nstep;
      w->dt = h.dt;
      w->df = h.df;
    }
    for (size_t i = 0; i < fr.nslc_count; ++i) {
      StepackNslc e;
      const size_t idx = fr.nslc_start + i;
      if (pread_exact(fd, sizeof(StepackHeader) + idx * sizeof(e), &e, sizeof(e)) != 0 ||
          e.file_index != idx || e.nsl_id == 0) {
        close(fd);
        FXH_LOG_ERROR("stepack_nslc_invalid", "path=\"%s\" file_index=%zu", fr.pack_path.c_str(),
                      idx);
        return false;
      }
      Row r;
      r.meta.gnsl_id = e.nsl_id;
      r.meta.stla = e.stla;
      r.meta.stlo = e.stlo;
      r.meta.network = fixed_read(e.network, sizeof(e.network));
      r.meta.station = fixed_read(e.station, sizeof(e.station));
      r.meta.location = fixed_read(e.location, sizeof(e.location));
      r.meta.component = fixed_read(e.component, sizeof(e.component));
      r.frag = fi;
      r.local = i;
      r.order = rows.size();
      rows.push_back(r);
    }
    close(fd);
    w->fragments.push_back(fr);
  }
  std::stable_sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) {
    return a.meta.gnsl_id < b.meta.gnsl_id;
  });
  for (const Row &r : rows) {
    w->specs.push_back(r.meta);
    w->locators.push_back({r.frag, r.local});
  }
  // every station carries the same number of channels
  size_t expected = 0;
  for (size_t i = 0; i < w->specs.size();) {
    size_t j = i;
    while (j < w->specs.size() && w->specs[j].gnsl_id == w->specs[i].gnsl_id) ++j;
    if (expected == 0) expected = j - i;
    if (j - i != expected) {
      FXH_LOG_ERROR("nslc_channel_group_incomplete", "timestamp=\"%s\" nsl_id=%u channels=%zu expected=%zu",
                    w->timestamp.c_str(), w->specs[i].gnsl_id, j - i, expected);
      return false;
    }
    i = j;
  }
  w->num_channels = expected;
  return expected > 0;
}

bool read_timestamp_payload(const TimestampWork &w, std::complex<float> *payload) {
  const size_t rows = w.specs.size();
  const size_t row_bytes = (size_t)w.nspec * 8;
  std::vector<int> fds(w.fragments.size(), -1);
  bool ok = true;
  for (size_t f = 0; f < w.fragments.size() && ok; ++f) {
    fds[f] = open(w.fragments[f].pack_path.c_str(), O_RDONLY);
    ok = fds[f] >= 0;
  }
  for (int s = 0; s < w.nstep && ok; ++s) {
    size_t i = 0;
    while (i < rows && ok) {
      // runs of rows contiguous in one fragment are read with one pread
      size_t j = i + 1;
      while (j < rows && w.locators[j].first == w.locators[i].first &&
             w.locators[j].second == w.locators[j - 1].second + 1)
        ++j;
      const StepackFragment &fr = w.fragments[w.locators[i].first];
      const size_t off = fr.payload_offset + (size_t)s * fr.pitch_step_bytes +
                         (fr.nslc_start + w.locators[i].second) * fr.nslc_step_bytes;
      char *dst = (char *)payload + ((size_t)s * rows + i) * row_bytes;
      ok = pread_exact(fds[w.locators[i].first], off, dst, (j - i) * row_bytes) == 0;
      i = j;
    }
  }
  for (int fd : fds)
    if (fd >= 0) close(fd);
  if (!ok) FXH_LOG_ERROR("stepack_payload_read_failed", "timestamp=\"%s\"", w.timestamp.c_str());
  return ok;
}

}  // namespace fxh
