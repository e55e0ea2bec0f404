// xc_fast job and task enumeration (host side).
//
// Restates native/xc/worker/scheduler.cu:43-152 and the allowed-path table of
// native/xc/io/path_table.c:9-191: stations are cut into blocks of
// block_size NSLC rows; every job pairs an anchor block with itself or with a
// later target block; inside a job the anchor rows are crossed with the
// target rows, a pair is kept when its canonical path id (lower gnsl id *
// 10000 + higher) is in the allowed table, and it is oriented src = lower
// gnsl id.  Rows of one station (several channels) give autocorrelation
// tasks, both orientations when the rows lie in different blocks.
#pragma once

#include "stepack.hpp"

#include <string>
#include <vector>

namespace fxh {

struct PathRecord {
  int path_id = 0, src_gnsl_id = 0, rec_gnsl_id = 0;
  float great_circle_deg = 0.f, distance_km = 0.f, azimuth_deg = 0.f, back_azimuth_deg = 0.f;
};

// 9-column TSV (path_id src_nsl_id rec_nsl_id src_nsl rec_nsl great_circle_deg
// distance_km azimuth_deg back_azimuth_deg) or the 8-column form with
// distance in column 5; header and '#' lines skipped, non-canonical rows
// dropped; sorted by path id.
bool load_path_table(const std::string &path, std::vector<PathRecord> *table);
int make_path_id(int src_gnsl, int rec_gnsl);  // -1 outside 1..9999
const PathRecord *find_path_canonical(const std::vector<PathRecord> &table, int a, int b);

struct XcJob {
  size_t anchor_block = 0, target_begin_block = 0, target_end_block = 0, block_size = 0;
  size_t file_count = 0, anchor_begin = 0, anchor_end = 0, target_begin = 0, target_end = 0;
};

struct XcTask {
  size_t src_meta = 0, rec_meta = 0;    // rows of the timestamp
  int src_local = 0, rec_local = 0;     // rows among the job's loaded rows
  int path_id = 0;
  bool is_autocorr = false;
  PathRecord path;
};

std::vector<XcJob> build_jobs(size_t file_count, size_t block_size);
// tasks in the reference order; `loaded` receives the job's loaded rows
// (anchor rows, then target rows when the target differs).
void build_tasks_for_job(const std::vector<SpecMeta> &specs, const std::vector<PathRecord> &table,
                         const XcJob &job, std::vector<XcTask> *tasks,
                         std::vector<size_t> *loaded);

}  // namespace fxh
