// StepPack v3: the spectra file sac2spec writes and xc_fast reads.
//
// File = 232-byte header + NSLC table (128 bytes per entry) + payload
// [nstep][batch_nslc][nspec] complex64, with a TSV sidecar of one row per
// timestamp run (native/sac2spec/io/stepack_writer.c:13-70, :433-535; read
// by native/xc/io/input.cu:429-866).
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace fxh {

constexpr uint32_t kStepackVersion = 3;
constexpr uint32_t kStepackLayout = 2;  // pitched [step][nslc][freq]

#pragma pack(push, 1)
struct StepackHeader {
  char magic[8];  // "FXCSTPK"
  uint32_t version, header_size, nslc_entry_size, layout;
  uint64_t batch_seq, start_group, group_count;
  uint32_t worker_id, nstep, nslc_count, nspec;
  float dt, df;
  uint64_t nslc_table_bytes, payload_offset, payload_bytes, pitch_step_bytes;
  char first_timestamp[64];
  unsigned char reserved[64];
};
struct StepackNslc {
  uint32_t file_index, nsl_id;
  float stla, stlo;
  char network[16], station[32], location[16], component[16];
  unsigned char reserved[32];
};
#pragma pack(pop)
static_assert(sizeof(StepackHeader) == 232, "StepPack header is 232 bytes");
static_assert(sizeof(StepackNslc) == 128, "StepPack NSLC entry is 128 bytes");

// One station-channel of a batch, in batch order.
struct StepackNode {
  std::string timestamp;  // as in sac_index.tsv (e.g. 2020.001.0000)
  std::string nsl_id;     // text id (TSV first/last columns)
  uint32_t gnsl_id = 0;
  float stla = 0.f, stlo = 0.f;
  std::string network, station, location, component;
};

// Writes <root>/w%03d.b%06zu.{stepack,tsv}; spectra [nstep][nodes][nspec].
int stepack_write_batch(const std::string &root, int worker_id, size_t batch_seq,
                        size_t start_group, size_t group_count,
                        const std::vector<StepackNode> &nodes, int nstep, int nspec, float dt,
                        float df, const std::complex<float> *spectra);

// ----------------------------------------------------------------- reading --
struct StepackFragment {
  std::string pack_path;
  size_t worker_id = 0, batch_seq = 0, start_group = 0, group_count = 0;
  size_t nstep = 0, nslc_start = 0, nslc_count = 0, batch_nslc_count = 0, nspec = 0;
  float dt = 0.f, df = 0.f;
  size_t payload_offset = 0, payload_bytes = 0, step_bytes = 0, pitch_step_bytes = 0,
         nslc_step_bytes = 0;
};

struct TimestampInput {
  std::string timestamp;  // normalised "YYYYMMDDTHH:MM"
  std::vector<StepackFragment> fragments;
};

// -I argument: a directory (its stepack/ subdirectory when present) of TSV
// sidecars, or one TSV.  Timestamps sorted; fragments by (batch, group, start,
// path).  Returns false (logged) on malformed input.
bool load_timestamp_inputs(const std::string &input_path, std::vector<TimestampInput> *out);

struct SpecMeta {
  uint32_t gnsl_id = 0;
  float stla = 0.f, stlo = 0.f;
  std::string network, station, location, component;
};

// One timestamp opened: NSLC rows stable-sorted by gnsl_id, the channel
// count per station, and the payload [nstep][rows][nspec] gathered into
// `payload` (caller-provided storage of nstep*rows*nspec complex values,
// e.g. pinned host memory; pass nullptr to size only).
struct TimestampWork {
  std::string timestamp;
  std::vector<SpecMeta> specs;
  int nstep = 0, nspec = 0;
  float dt = 0.f, df = 0.f;
  size_t num_channels = 0;
  // (fragment index, row within the fragment) of every sorted row
  std::vector<std::pair<size_t, size_t>> locators;
  std::vector<StepackFragment> fragments;
};

bool open_timestamp_work(const TimestampInput &in, TimestampWork *w);
bool read_timestamp_payload(const TimestampWork &w, std::complex<float> *payload);

}  // namespace fxh
