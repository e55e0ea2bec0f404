#include "fxc_host.hpp"

#include <dirent.h>
#include <errno.h>
#include <fcntl.h>
#include <limits.h>
#include <strings.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>

namespace fxh {

// ---------------------------------------------------------------- logging --
namespace {
LogLevel g_level = kLogInfo;
bool g_level_set = false;

const char *level_name(LogLevel l) {
  switch (l) {
    case kLogError: return "ERROR";
    case kLogWarn: return "WARN";
    case kLogInfo: return "INFO";
    case kLogDebug: return "DEBUG";
  }
  return "LOG";
}
}  // namespace

LogLevel log_level() {
  if (g_level_set) return g_level;
  g_level_set = true;
  const char *env = getenv("FASTXC_LOG_LEVEL");
  if (!env || !*env) env = getenv("XC_LOG_LEVEL");
  if (!env || !*env) env = getenv("SAC2SPEC_LOG_LEVEL");
  if (!env || !*env) return g_level;
  if (!strcmp(env, "0") || !strcasecmp(env, "ERROR")) g_level = kLogError;
  else if (!strcmp(env, "1") || !strcasecmp(env, "WARN") || !strcasecmp(env, "WARNING"))
    g_level = kLogWarn;
  else if (!strcmp(env, "2") || !strcasecmp(env, "INFO")) g_level = kLogInfo;
  else if (!strcmp(env, "3") || !strcasecmp(env, "DEBUG")) g_level = kLogDebug;
  return g_level;
}

void log_event(LogLevel level, const char *event, const char *file, int line, const char *fmt,
               ...) {
  if (level > log_level()) return;
  const char *base = file ? strrchr(file, '/') : nullptr;
  flockfile(stderr);
  fprintf(stderr, "%s event=%s file=%s line=%d ", level_name(level), event ? event : "unknown",
          base ? base + 1 : (file ? file : "unknown"), line);
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fputc('\n', stderr);
  funlockfile(stderr);
}

// --------------------------------------------------------------- progress --
void Progress::init(const char *path) {
  rows_.clear();
  enabled_ = path && *path && strcmp(path, "NONE") != 0;
  path_ = enabled_ ? path : "";
}

void Progress::set_rows(const std::vector<ProgressRow> &rows) {
  if (!enabled_) return;
  rows_ = rows;
  write();
}

ProgressRow *Progress::find(const std::string &task) {
  for (auto &r : rows_)
    if (r.task == task) return &r;
  return nullptr;
}

void Progress::update(const std::string &task, const std::string &status, size_t completed,
                      size_t total, const std::string &unit, const std::string &detail) {
  if (!enabled_) return;
  ProgressRow *r = find(task);
  if (!r) {
    rows_.push_back({task, status, completed, total, unit, detail});
  } else {
    r->status = status;
    r->completed = completed;
    r->total = total;
    r->unit = unit;
    r->detail = detail;
  }
  write();
}

void Progress::add(const std::string &task, size_t delta, const std::string &detail) {
  if (!enabled_) return;
  ProgressRow *r = find(task);
  if (!r) return;
  r->completed = std::min(r->completed + delta, r->total);
  r->status = (r->total > 0 && r->completed >= r->total) ? "DONE" : "RUNNING";
  if (!detail.empty()) r->detail = detail;
  write();
}

void Progress::finish(const std::string &status, bool complete) {
  if (!enabled_) return;
  for (auto &r : rows_) {
    r.status = status;
    if (complete) r.completed = r.total;
  }
  write();
}

void Progress::write() {
  const std::string tmp = path_ + ".tmp";
  FILE *fp = fopen(tmp.c_str(), "w");
  if (!fp) return;
  fputs("task\tstatus\tcompleted\ttotal\tunit\tdetail\n", fp);
  for (const auto &r : rows_)
    fprintf(fp, "%s\t%s\t%zu\t%zu\t%s\t%s\n", clean_tsv_field(r.task).c_str(),
            clean_tsv_field(r.status).c_str(), r.completed, r.total,
            clean_tsv_field(r.unit).c_str(), clean_tsv_field(r.detail).c_str());
  if (fclose(fp) == 0) rename(tmp.c_str(), path_.c_str());
}

// -------------------------------------------------------------------- TSV --
std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && isspace((unsigned char)s[a])) ++a;
  while (b > a && isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::vector<std::string> split_tab(const std::string &s) {
  std::vector<std::string> out;
  size_t start = 0;
  for (;;) {
    const size_t t = s.find('\t', start);
    if (t == std::string::npos) {
      out.push_back(s.substr(start));
      break;
    }
    out.push_back(s.substr(start, t - start));
    start = t + 1;
  }
  return out;
}

int column_index(const std::vector<std::string> &header, const char *name) {
  for (size_t i = 0; i < header.size(); ++i)
    if (trim(header[i]) == name) return (int)i;
  return -1;
}

std::string clean_tsv_field(const std::string &s) {
  std::string o = s;
  for (char &c : o)
    if (c == '\t' || c == '\n' || c == '\r') c = ' ';
  return o;
}

bool parse_size_strict(const std::string &s, size_t *out) {
  const std::string t = trim(s);
  if (t.empty() || t[0] == '-') return false;
  errno = 0;
  char *end = nullptr;
  const unsigned long long v = strtoull(t.c_str(), &end, 10);
  if (errno || !end || *end) return false;
  *out = (size_t)v;
  return true;
}

bool parse_float_strict(const std::string &s, float *out) {
  const std::string t = trim(s);
  if (t.empty()) return false;
  errno = 0;
  char *end = nullptr;
  const float v = strtof(t.c_str(), &end);
  if (errno || !end || *end || !std::isfinite(v)) return false;
  *out = v;
  return true;
}

// -------------------------------------------------------------------- fs --
bool is_dir(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool is_file(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

int mkdir_p(const std::string &p) {
  if (p.empty()) return -1;
  std::string cur;
  size_t i = 0;
  if (p[0] == '/') {
    cur = "/";
    i = 1;
  }
  while (i <= p.size()) {
    const size_t j = p.find('/', i);
    const std::string part = p.substr(i, j == std::string::npos ? std::string::npos : j - i);
    if (!part.empty()) {
      cur += part;
      if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return -1;
      if (!is_dir(cur)) return -1;
      cur += "/";
    }
    if (j == std::string::npos) break;
    i = j + 1;
  }
  return 0;
}

std::string dirname_of(const std::string &p) {
  const size_t s = p.find_last_of('/');
  if (s == std::string::npos) return ".";
  if (s == 0) return "/";
  return p.substr(0, s);
}

std::string path_join(const std::string &a, const std::string &b) {
  if (b.empty()) return a;
  if (!b.empty() && b[0] == '/') return b;
  if (a.empty()) return b;
  return a.back() == '/' ? a + b : a + "/" + b;
}

std::string absolute_path(const std::string &p) {
  if (!p.empty() && p[0] == '/') return p;
  char buf[PATH_MAX];
  if (!getcwd(buf, sizeof(buf))) return p;
  return path_join(buf, p);
}

std::vector<std::string> list_files_with_suffix(const std::string &dir, const char *suffix) {
  std::vector<std::string> out;
  DIR *d = opendir(dir.c_str());
  if (!d) return out;
  const size_t sl = strlen(suffix);
  while (struct dirent *e = readdir(d)) {
    const std::string name = e->d_name;
    if (name.size() >= sl && name.compare(name.size() - sl, sl, suffix) == 0) {
      const std::string full = path_join(dir, name);
      if (is_file(full)) out.push_back(full);
    }
  }
  closedir(d);
  std::sort(out.begin(), out.end());
  return out;
}

bool dir_has_entries(const std::string &dir) {
  DIR *d = opendir(dir.c_str());
  if (!d) return false;
  bool any = false;
  while (struct dirent *e = readdir(d)) {
    if (strcmp(e->d_name, ".") && strcmp(e->d_name, "..")) {
      any = true;
      break;
    }
  }
  closedir(d);
  return any;
}

int write_marker(const std::string &path) {
  FILE *fp = fopen(path.c_str(), "w");
  if (!fp) return -1;
  const int w = fprintf(fp, "DONE\n");
  return (fclose(fp) == 0 && w > 0) ? 0 : -1;
}

int pread_exact(int fd, size_t offset, void *dst, size_t bytes) {
  size_t done = 0;
  char *out = (char *)dst;
  while (done < bytes) {
    const size_t chunk = std::min(bytes - done, (size_t)1 << 30);
    const ssize_t n = pread(fd, out + done, chunk, (off_t)(offset + done));
    if (n <= 0) return -1;
    done += (size_t)n;
  }
  return 0;
}

// ------------------------------------------------------------------- SAC --
SacHeader sac_null() {
  SacHeader h;
  float *f = (float *)&h;
  for (int i = 0; i < 70; ++i) f[i] = -12345.f;
  int32_t *n = (int32_t *)(f + 70);
  for (int i = 0; i < 40; ++i) n[i] = -12345;
  h.nvhdr = 6;
  // 23 text fields: kevnm (the second) is 16 bytes, the other 22 are 8
  char *c = (char *)(n + 40);
  for (int i = 0; i < 23; ++i) {
    const int len = i == 1 ? 16 : 8;
    memset(c, ' ', len);
    memcpy(c, "-12345", 6);
    c += len;
  }
  return h;
}

void sac_text(char *dst, size_t size, const std::string &s) {
  memset(dst, ' ', size);
  memcpy(dst, s.data(), std::min(size, s.size()));
}

std::string sac_string(const char *src, size_t size) {
  size_t n = 0;
  while (n < size && src[n] != '\0') ++n;
  while (n > 0 && (src[n - 1] == ' ' || src[n - 1] == '\0')) --n;
  return std::string(src, n);
}

int read_sac_header(const std::string &path, SacHeader *hd) {
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) {
    FXH_LOG_ERROR("sac_open_failed", "path=\"%s\" error=\"%s\"", path.c_str(), strerror(errno));
    return -1;
  }
  const size_t r = fread(hd, sizeof(*hd), 1, fp);
  fclose(fp);
  if (r != 1) {
    FXH_LOG_ERROR("sac_header_read_failed", "path=\"%s\"", path.c_str());
    return -1;
  }
  return 0;
}

int read_sac(const std::string &path, SacHeader *hd, std::vector<float> *data,
             size_t max_samples) {
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) {
    FXH_LOG_ERROR("sac_open_failed", "path=\"%s\" error=\"%s\"", path.c_str(), strerror(errno));
    return -1;
  }
  if (fread(hd, sizeof(*hd), 1, fp) != 1) {
    fclose(fp);
    FXH_LOG_ERROR("sac_header_read_failed", "path=\"%s\"", path.c_str());
    return -1;
  }
  data->assign(max_samples, 0.f);
  const size_t want = hd->npts > 0 ? std::min((size_t)hd->npts, max_samples) : 0;
  const size_t got = want ? fread(data->data(), sizeof(float), want, fp) : 0;
  fclose(fp);
  if (got != want) {
    FXH_LOG_ERROR("sac_data_read_failed", "path=\"%s\" expected=%zu read=%zu", path.c_str(), want,
                  got);
    return -1;
  }
  return 0;
}

int write_sac(const std::string &path, const SacHeader &hd, const float *data, size_t n) {
  FILE *fp = fopen(path.c_str(), "wb");
  if (!fp) return -1;
  const bool ok = fwrite(&hd, sizeof(hd), 1, fp) == 1 && fwrite(data, sizeof(float), n, fp) == n;
  return (fclose(fp) == 0 && ok) ? 0 : -1;
}

// ------------------------------------------------------------- timestamps --
namespace {
bool leap(int y) { return (y % 4 == 0 && y % 100 != 0) || y % 400 == 0; }

bool date_from_jday(int year, int jday, int *month, int *day) {
  static const int mdays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  if (year <= 0 || jday < 1 || jday > (leap(year) ? 366 : 365)) return false;
  int m = 0, d = jday;
  while (m < 12) {
    const int md = mdays[m] + (m == 1 && leap(year) ? 1 : 0);
    if (d <= md) break;
    d -= md;
    ++m;
  }
  *month = m + 1;
  *day = d;
  return true;
}

int jday_from_ymd(int y, int m, int d) {
  static const int mdays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  if (y <= 0 || m < 1 || m > 12) return 0;
  const int md = mdays[m - 1] + (m == 2 && leap(y) ? 1 : 0);
  if (d < 1 || d > md) return 0;
  int j = d;
  for (int i = 0; i < m - 1; ++i) j += mdays[i] + (i == 1 && leap(y) ? 1 : 0);
  return j;
}

bool clock_ok(int h, int mi, int s) { return h >= 0 && h <= 23 && mi >= 0 && mi <= 59 && s >= 0 && s <= 59; }

std::string fmt_xc(int y, int mo, int d, int h, int mi) {
  char out[40];
  snprintf(out, sizeof(out), "%04d%02d%02dT%02d:%02d", y, mo, d, h, mi);
  return out;
}

bool all_digits(const std::string &s, size_t pos, size_t len, int *v) {
  if (pos + len > s.size()) return false;
  int x = 0;
  for (size_t i = pos; i < pos + len; ++i) {
    if (!isdigit((unsigned char)s[i])) return false;
    x = x * 10 + (s[i] - '0');
  }
  *v = x;
  return true;
}
}  // namespace

std::string normalize_timestamp_for_xc(const std::string &text) {
  const std::string raw = trim(text);
  int y = 0, mo = 0, d = 0, j = 0, h = 0, mi = 0, s = 0, n = 0;
  const int len = (int)raw.size();
  // year.jday.HHMM and year.jday.HH:MM
  if (sscanf(raw.c_str(), "%d.%d.%2d%2d%n", &y, &j, &h, &mi, &n) == 4 && n == len &&
      date_from_jday(y, j, &mo, &d) && clock_ok(h, mi, 0))
    return fmt_xc(y, mo, d, h, mi);
  n = 0;
  if (sscanf(raw.c_str(), "%d.%d.%d:%d%n", &y, &j, &h, &mi, &n) == 4 && n == len &&
      date_from_jday(y, j, &mo, &d) && clock_ok(h, mi, 0))
    return fmt_xc(y, mo, d, h, mi);
  // calendar spellings with or without seconds, 'T' or ' ' separators (the
  // clock is validated, the calendar date is passed through)
  static const char *const six[] = {"%4d-%2d-%2dT%2d:%2d:%2d%n", "%4d-%2d-%2d %2d:%2d:%2d%n",
                                    "%4d%2d%2dT%2d:%2d:%2d%n"};
  static const char *const five[] = {"%4d-%2d-%2dT%2d:%2d%n", "%4d-%2d-%2d %2d:%2d%n",
                                     "%4d%2d%2dT%2d:%2d%n", "%4d%2d%2dT%2d%2d%n",
                                     "%4d%2d%2d%2d%2d%n"};
  for (const char *f : six) {
    n = 0;
    if (sscanf(raw.c_str(), f, &y, &mo, &d, &h, &mi, &s, &n) == 6 && n == len &&
        clock_ok(h, mi, s))
      return fmt_xc(y, mo, d, h, mi);
  }
  for (const char *f : five) {
    n = 0;
    if (sscanf(raw.c_str(), f, &y, &mo, &d, &h, &mi, &n) == 5 && n == len && clock_ok(h, mi, 0))
      return fmt_xc(y, mo, d, h, mi);
  }
  FXH_LOG_ERROR("stepack_timestamp_unsupported",
                "timestamp=\"%s\" expected=\"YYYY.JJJ.HHMM or YYYYMMDDTHH:MM\"", text.c_str());
  return std::string();
}

bool parse_timestamp_text(const std::string &ts, TimeData *out) {
  *out = TimeData();
  int y = 0, j = 0, hhmm = 0, used = 0;
  if (sscanf(ts.c_str(), "%d.%d.%d%n", &y, &j, &hhmm, &used) == 3 && used == (int)ts.size()) {
    const int h = hhmm / 100, mi = hhmm % 100;
    if (y > 0 && j >= 1 && j <= (leap(y) ? 366 : 365) && clock_ok(h, mi, 0)) {
      *out = {y, j, h, mi};
      return true;
    }
    return false;
  }
  int mo = 0, d = 0, h = 0, mi = 0;
  size_t tp = 0;
  if (ts.size() >= 13 && (ts[8] == 'T' || ts[8] == 't')) {
    if (!all_digits(ts, 0, 4, &y) || !all_digits(ts, 4, 2, &mo) || !all_digits(ts, 6, 2, &d))
      return false;
    tp = 9;
  } else if (ts.size() >= 15 && ts[4] == '-' && ts[7] == '-' && (ts[10] == 'T' || ts[10] == 't')) {
    if (!all_digits(ts, 0, 4, &y) || !all_digits(ts, 5, 2, &mo) || !all_digits(ts, 8, 2, &d))
      return false;
    tp = 11;
  } else {
    return false;
  }
  if (tp + 5 <= ts.size() && ts[tp + 2] == ':') {
    if (!all_digits(ts, tp, 2, &h) || !all_digits(ts, tp + 3, 2, &mi)) return false;
  } else if (tp + 4 <= ts.size()) {
    if (!all_digits(ts, tp, 2, &h) || !all_digits(ts, tp + 2, 2, &mi)) return false;
  } else {
    return false;
  }
  const int jd = jday_from_ymd(y, mo, d);
  if (!jd || !clock_ok(h, mi, 0)) return false;
  *out = {y, jd, h, mi};
  return true;
}

}  // namespace fxh
