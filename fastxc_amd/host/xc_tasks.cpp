#include "xc_tasks.hpp"

#include "fxc_host.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>

namespace fxh {

int make_path_id(int src, int rec) {
  if (src <= 0 || src > 9999 || rec <= 0 || rec > 9999) return -1;
  return src * 10000 + rec;
}

bool load_path_table(const std::string &path, std::vector<PathRecord> *table) {
  std::ifstream in(path);
  if (!in) {
    FXH_LOG_ERROR("allowed_path_table_open_failed", "path=\"%s\"", path.c_str());
    return false;
  }
  table->clear();
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#' || line.compare(0, 7, "path_id") == 0) continue;
    std::vector<std::string> f;
    for (const auto &t : split_tab(line)) {
      const std::string s = trim(t);
      if (!s.empty()) f.push_back(s);
    }
    if (f.size() < 8) continue;
    PathRecord r;
    r.path_id = atoi(f[0].c_str());
    r.src_gnsl_id = atoi(f[1].c_str());
    r.rec_gnsl_id = atoi(f[2].c_str());
    if (f.size() >= 9) {
      r.great_circle_deg = (float)atof(f[5].c_str());
      r.distance_km = (float)atof(f[6].c_str());
      r.azimuth_deg = (float)atof(f[7].c_str());
      r.back_azimuth_deg = (float)atof(f[8].c_str());
    } else {
      r.great_circle_deg = (float)(atof(f[5].c_str()) / 111.195);
      r.distance_km = (float)atof(f[5].c_str());
      r.azimuth_deg = (float)atof(f[6].c_str());
      r.back_azimuth_deg = (float)atof(f[7].c_str());
    }
    if (r.path_id <= 0 || r.path_id != make_path_id(r.src_gnsl_id, r.rec_gnsl_id) ||
        r.src_gnsl_id > r.rec_gnsl_id) {
      FXH_LOG_WARN("allowed_path_row_skipped",
                   "path=\"%s\" path_id=%d src=%d rec=%d reason=malformed_or_noncanonical",
                   path.c_str(), r.path_id, r.src_gnsl_id, r.rec_gnsl_id);
      continue;
    }
    table->push_back(r);
  }
  std::stable_sort(table->begin(), table->end(),
                   [](const PathRecord &a, const PathRecord &b) { return a.path_id < b.path_id; });
  FXH_LOG_INFO("allowed_path_table_loaded", "path=\"%s\" records=%zu", path.c_str(),
               table->size());
  return true;
}

const PathRecord *find_path_canonical(const std::vector<PathRecord> &table, int a, int b) {
  const int id = make_path_id(std::min(a, b), std::max(a, b));
  if (id <= 0) return nullptr;
  auto it = std::lower_bound(table.begin(), table.end(), id,
                             [](const PathRecord &r, int v) { return r.path_id < v; });
  return (it != table.end() && it->path_id == id) ? &*it : nullptr;
}

std::vector<XcJob> build_jobs(size_t file_count, size_t block_size) {
  std::vector<XcJob> jobs;
  if (file_count == 0 || block_size == 0) return jobs;
  const size_t nblk = (file_count + block_size - 1) / block_size;
  auto fill = [&](size_t anchor, size_t tb, size_t te) {
    XcJob j;
    j.anchor_block = anchor;
    j.target_begin_block = tb;
    j.target_end_block = te;
    j.block_size = block_size;
    j.file_count = file_count;
    j.anchor_begin = anchor * block_size;
    j.anchor_end = std::min(file_count, j.anchor_begin + block_size);
    j.target_begin = tb * block_size;
    j.target_end = std::min(file_count, te * block_size);
    return j;
  };
  for (size_t a = 0; a < nblk; ++a) {
    jobs.push_back(fill(a, a, a + 1));
    for (size_t t = a + 1; t < nblk; ++t) jobs.push_back(fill(a, t, t + 1));
  }
  return jobs;
}

void build_tasks_for_job(const std::vector<SpecMeta> &specs, const std::vector<PathRecord> &table,
                         const XcJob &job, std::vector<XcTask> *tasks,
                         std::vector<size_t> *loaded) {
  tasks->clear();
  loaded->clear();
  for (size_t i = job.anchor_begin; i < job.anchor_end; ++i) loaded->push_back(i);
  if (job.target_begin != job.anchor_begin)
    for (size_t i = job.target_begin; i < job.target_end; ++i) loaded->push_back(i);
  auto local = [&](size_t m) -> int {
    if (m >= job.anchor_begin && m < job.anchor_end) return (int)(m - job.anchor_begin);
    return (int)((job.anchor_end - job.anchor_begin) + (m - job.target_begin));
  };
  auto push = [&](size_t s, size_t r, const PathRecord *rec, bool autoc) {
    XcTask t;
    t.src_meta = s;
    t.rec_meta = r;
    t.src_local = local(s);
    t.rec_local = local(r);
    t.path_id = rec->path_id;
    t.path = *rec;
    t.is_autocorr = autoc;
    tasks->push_back(t);
  };
  for (size_t tb = job.target_begin_block; tb < job.target_end_block; ++tb) {
    size_t begin = tb * job.block_size;
    size_t end = std::min(job.file_count, begin + job.block_size);
    const bool same = tb == job.anchor_block;
    if (same) {
      begin = job.anchor_begin;
      end = job.anchor_end;
    }
    for (size_t a = job.anchor_begin; a < job.anchor_end; ++a)
      for (size_t b = begin; b < end; ++b) {
        const int ga = (int)specs[a].gnsl_id, gb = (int)specs[b].gnsl_id;
        const bool autoc = ga == gb;
        if (same && !autoc && b <= a) continue;
        const PathRecord *rec = find_path_canonical(table, ga, gb);
        if (!rec) continue;
        if (autoc) {
          push(a, b, rec, true);
          if (!same) push(b, a, rec, true);
        } else if (ga < gb) {
          push(a, b, rec, false);
        } else {
          push(b, a, rec, false);
        }
      }
  }
}

}  // namespace fxh
