// sourcepack_index: per-timestamp SourcePack indexes over xc_fast's xcpack
// output, the batch form of the indexer xc_fast runs with --sourcepack-dir.
//
//   sourcepack_index -I <xc output root|xcpack dir> -O <sourcepack dir>
//                    [--sort | --keep-order] [--timestamp TS] [--progress FILE]
//
// Options and the summary line follow `fastxc sourcepack` (fastxc/cli.py:
// 127-141, 474-491); the build is build_sourcepack / build_sourcepack_timestamp
// (fastxc/operators/sourcepack/builder.py:163-264). Exit status 0 on success,
// 1 on any error (logged).
#include "fxc_host.hpp"
#include "sourcepack.hpp"

#include <cstdio>
#include <cstring>
#include <string>

using namespace fxh;

namespace {

void usage() {
  fprintf(stderr,
          "\nUsage:\n"
          "  sourcepack_index -I <xc_output_root|xcpack_dir> -O <sourcepack_dir> [options]\n\n"
          "  --sort               Sort each index by path, component slot, source,\n"
          "                       receiver (default)\n"
          "  --keep-order         Keep the xcpack TSV order\n"
          "  --timestamp <ts>     Build this timestamp only (no global _SUCCESS)\n"
          "  --progress <file>    Write progress sidecar TSV\n\n");
}

}  // namespace

int main(int argc, char **argv) {
  std::string input, output, timestamp, progress_path;
  bool sort = true;
  for (int i = 1; i < argc; ++i) {
    const std::string o = argv[i];
    auto value = [&]() -> const char * {
      if (i + 1 >= argc) {
        FXH_LOG_ERROR("argument_value_missing", "option=\"%s\"", o.c_str());
        return nullptr;
      }
      return argv[++i];
    };
    const char *v = nullptr;
    if (o == "-h" || o == "--help") {
      usage();
      return 0;
    } else if (o == "-I" || o == "--input") {
      if (!(v = value())) return 1;
      input = v;
    } else if (o == "-O" || o == "--output") {
      if (!(v = value())) return 1;
      output = v;
    } else if (o == "--timestamp") {
      if (!(v = value())) return 1;
      timestamp = v;
    } else if (o == "--progress" || o == "--progress-file") {
      if (!(v = value())) return 1;
      progress_path = v;
    } else if (o == "--sort") {
      sort = true;
    } else if (o == "--keep-order") {
      sort = false;
    } else {
      FXH_LOG_ERROR("argument_unknown", "option=\"%s\"", o.c_str());
      usage();
      return 1;
    }
  }
  if (input.empty() || output.empty()) {
    usage();
    return 1;
  }
  const std::string xcpack = xcpack_dir_of(input);
  if (!is_dir(xcpack)) {
    FXH_LOG_ERROR("xcpack_dir_missing", "input=\"%s\" candidate=\"%s\"", input.c_str(),
                  xcpack.c_str());
    return 1;
  }
  const std::string out = absolute_path(output);
  SourcePackResult res;
  if (!timestamp.empty()) {
    if (!sourcepack_build_timestamp(xcpack, out, timestamp, sort, &res)) return 1;
  } else {
    Progress prog;
    if (!progress_path.empty() && progress_path != "NONE") mkdir_p(dirname_of(progress_path));
    prog.init(progress_path.c_str());
    prog.update("overall", "RUNNING", 0, 0, "timestamps", "source index build");
    if (!sourcepack_build_all(xcpack, out, sort, &res)) {
      prog.update("overall", "FAILED", res.index_paths.size(), 0, "timestamps",
                  "source index build failed");
      return 1;
    }
    prog.update("overall", "DONE", res.index_paths.size(), res.index_paths.size(), "timestamps",
                "source index ready");
  }
  printf("Built %zu timestamp source index(es), %lld source(s), %lld record(s).\n",
         res.index_paths.size(), res.sources, res.records);
  return 0;
}
