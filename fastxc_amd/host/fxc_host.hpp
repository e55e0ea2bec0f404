// Host-side pieces shared by the stage executables (sac2spec, xc_fast,
// ncf_pws, ncf_tfpws): structured logging, the progress sidecar, TSV and
// file-system helpers, the SAC header, and timestamp text handling.  The
// byte formats and text formats follow the reference so the Python stage
// contract reads these programs' outputs unchanged.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace fxh {

// ---------------------------------------------------------------- logging --
// One line per event on stderr: "LEVEL event=<name> file=<f> line=<n> k=v ...",
// level from FASTXC_LOG_LEVEL, else XC_LOG_LEVEL, else SAC2SPEC_LOG_LEVEL
// (ERROR|WARN|INFO|DEBUG or 0..3; default INFO), as native/xc/logger.c.
enum LogLevel { kLogError = 0, kLogWarn = 1, kLogInfo = 2, kLogDebug = 3 };
void log_event(LogLevel level, const char *event, const char *file, int line, const char *fmt,
               ...) __attribute__((format(printf, 5, 6)));
LogLevel log_level();

#define FXH_LOG_ERROR(ev, ...) ::fxh::log_event(::fxh::kLogError, ev, __FILE__, __LINE__, __VA_ARGS__)
#define FXH_LOG_WARN(ev, ...) ::fxh::log_event(::fxh::kLogWarn, ev, __FILE__, __LINE__, __VA_ARGS__)
#define FXH_LOG_INFO(ev, ...) ::fxh::log_event(::fxh::kLogInfo, ev, __FILE__, __LINE__, __VA_ARGS__)
#define FXH_LOG_DEBUG(ev, ...) ::fxh::log_event(::fxh::kLogDebug, ev, __FILE__, __LINE__, __VA_ARGS__)

// --------------------------------------------------------------- progress --
// Sidecar TSV "task status completed total unit detail", rewritten whole
// through <path>.tmp + rename on every change (native/progress_sidecar.hpp).
// A null path, "" or "NONE" disables it.
struct ProgressRow {
  std::string task, status;
  size_t completed = 0, total = 0;
  std::string unit, detail;
};

class Progress {
 public:
  void init(const char *path);
  bool enabled() const { return enabled_; }
  void set_rows(const std::vector<ProgressRow> &rows);
  void update(const std::string &task, const std::string &status, size_t completed, size_t total,
              const std::string &unit, const std::string &detail);
  void add(const std::string &task, size_t delta, const std::string &detail);
  void finish(const std::string &status, bool complete);

 private:
  void write();
  ProgressRow *find(const std::string &task);
  std::string path_;
  bool enabled_ = false;
  std::vector<ProgressRow> rows_;
};

// -------------------------------------------------------------------- TSV --
std::string trim(const std::string &s);
std::vector<std::string> split_tab(const std::string &s);
int column_index(const std::vector<std::string> &header, const char *name);
std::string clean_tsv_field(const std::string &s);  // tabs/newlines -> spaces
bool parse_size_strict(const std::string &s, size_t *out);
bool parse_float_strict(const std::string &s, float *out);

// -------------------------------------------------------------------- fs --
bool is_dir(const std::string &p);
bool is_file(const std::string &p);
int mkdir_p(const std::string &p);
std::string dirname_of(const std::string &p);
std::string path_join(const std::string &a, const std::string &b);
std::string absolute_path(const std::string &p);
std::vector<std::string> list_files_with_suffix(const std::string &dir, const char *suffix);
int write_marker(const std::string &path);  // "DONE\n"
bool dir_has_entries(const std::string &dir);
int pread_exact(int fd, size_t offset, void *dst, size_t bytes);

// ------------------------------------------------------------------- SAC --
// The 632-byte SAC header: 70 floats, 40 ints (35 integer/enumerated + 5
// logical), 23 character fields (kevnm 16 bytes, the other 22 8 bytes).
#pragma pack(push, 1)
struct SacHeader {
  float delta, depmin, depmax, scale, odelta;
  float b, e, o, a, internal1;
  float t0, t1, t2, t3, t4, t5, t6, t7, t8, t9;
  float f, resp0, resp1, resp2, resp3, resp4, resp5, resp6, resp7, resp8, resp9;
  float stla, stlo, stel, stdp, evla, evlo, evel, evdp, mag;
  float user0, user1, user2, user3, user4, user5, user6, user7, user8, user9;
  float dist, az, baz, gcarc, internal2, internal3, depmen, cmpaz, cmpinc;
  float xminimum, xmaximum, yminimum, ymaximum;
  float unused_f[7];
  int32_t nzyear, nzjday, nzhour, nzmin, nzsec, nzmsec, nvhdr, norid, nevid, npts;
  int32_t internal_i[2], nxsize_unused[3];
  int32_t iftype, idep, iztype, unused16, iinst, istreg, ievreg, ievtyp, iqual, isynth;
  int32_t unused_i[10];
  int32_t leven, lpspol, lovrok, lcalda, unused27;
  char kstnm[8], kevnm[16], khole[8], ko[8], ka[8];
  char kt[10][8];
  char kf[8], kuser0[8], kuser1[8], kuser2[8], kcmpnm[8], knetwk[8], kdatrd[8], kinst[8];
};
#pragma pack(pop)
static_assert(sizeof(SacHeader) == 632, "SAC header is 632 bytes");

// Every float and int -12345, every string "-12345" padded with blanks,
// header version 6 (native/sac2spec/io/sacio.c:18-52).
SacHeader sac_null();
void sac_text(char *dst, size_t size, const std::string &s);  // blank-padded copy
std::string sac_string(const char *src, size_t size);          // trailing blanks/NULs cut

// Reads a SAC file's header and up to max_samples samples (zero-padded to
// max_samples).  Returns 0, or -1 with a logged reason.
int read_sac(const std::string &path, SacHeader *hd, std::vector<float> *data, size_t max_samples);
int read_sac_header(const std::string &path, SacHeader *hd);
int write_sac(const std::string &path, const SacHeader &hd, const float *data, size_t n);

// ------------------------------------------------------------- timestamps --
// "YYYY.JJJ.HHMM" (and the other accepted spellings) -> "YYYYMMDDTHH:MM",
// or "" when unsupported (native/xc/io/input.cu:106-224).
std::string normalize_timestamp_for_xc(const std::string &text);
struct TimeData {
  int year = 0, day_of_year = 0, hour = 0, minute = 0;
};
// Legacy "YYYY.JJJ.HHMM" or modern "YYYYMMDDTHH:MM"/"YYYY-MM-DDTHH:MM".
bool parse_timestamp_text(const std::string &ts, TimeData *out);

}  // namespace fxh
