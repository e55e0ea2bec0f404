// SourcePack indexes over xcpack output: the records of one timestamp
// grouped for the stacks (path, component slot, source, receiver), pointing
// at the xcpack bytes (no NCF bytes are copied).
//
// Restates build_sourcepack / build_sourcepack_timestamp /
// AsyncSourcePackMaterializer.finish (fastxc/operators/sourcepack/builder.py:
// 163-340, 38-118) and read_xcpack_tsv / station_key
// (fastxc/io/sourcepack.py:120-200): the same row columns
// (SOURCEPACK_INDEX_HEADER), sort key, csv line ends ("\r\n"), directory
// names, temporary-directory rename and _SUCCESS texts.
#pragma once

#include "fxc_host.hpp"

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

namespace fxh {

struct SourcePackResult {
  std::vector<std::string> index_paths;
  long long records = 0, sources = 0;
  uint64_t bytes_indexed = 0;
};

// "<root>" when its name is xcpack, else "<root>/xcpack"; symlinks resolved.
std::string xcpack_dir_of(const std::string &root);
// Characters outside [0-9A-Za-z._:-] become "-", "-" trimmed; "item" if empty.
std::string sourcepack_safe_name(const std::string &text);
// One <out>/<safe ts>/{sourcepack_index.tsv,_SUCCESS} from the xcpack TSVs of
// the timestamp (names starting with it, else every TSV); false on error.
// Rows sorted by (path_id, component slot, source, receiver, sequence), or
// in sequence (read) order when sort_within_source is false.
bool sourcepack_build_timestamp(const std::string &xcpack_dir, const std::string &out_dir,
                                const std::string &timestamp, bool sort_within_source,
                                SourcePackResult *result);
// Every timestamp that has records, a .done marker or a row in the
// workspace's manifest/timestamp_index.tsv.
bool sourcepack_build_all(const std::string &xcpack_dir, const std::string &out_dir,
                          bool sort_within_source, SourcePackResult *result);
// <out>/_SUCCESS: timestamps (index files present), records, sources, bytes.
bool sourcepack_write_success(const std::string &out_dir, const SourcePackResult &total);
// Timestamps the progress total counts: the workspace manifest's, else the
// .done markers present now (builder.py:355-359).
size_t sourcepack_expected_count(const std::string &xcpack_dir);

// Builds each timestamp's index on a thread of its own as the producer marks
// the timestamp done, so indexing overlaps the next timestamp's correlation
// (AsyncSourcePackMaterializer, builder.py:26-160, without the polling: the
// producer hands over each timestamp when it writes <ts>.done).
class SourcePackMaterializer {
 public:
  SourcePackMaterializer(const std::string &xc_output_root, const std::string &sourcepack_dir,
                         bool sort_within_source, const std::string &progress_file);
  // Without a finish call: finish(false, ...) (the producer failed).
  ~SourcePackMaterializer();
  void start();
  void submit(const std::string &timestamp);
  // Joins the thread, then builds every <ts>.done the thread has not built.
  // With mark_success, writes <out>/_SUCCESS; else removes a stale one.
  // False when any build failed.
  bool finish(bool mark_success, const std::string &status_message);
  const SourcePackResult &result() const { return total_; }

 private:
  void run();
  bool build(const std::string &timestamp);

  std::string xcpack_dir_, out_dir_;
  bool sort_;
  Progress progress_;
  size_t expected_ = 0;
  std::thread thread_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::string> queue_;
  std::set<std::string> done_;
  bool stop_ = false, failed_ = false, started_ = false, finished_ = false;
  std::string first_error_;
  SourcePackResult total_;
};

}  // namespace fxh
