// ckpt.hip - the checkpoint directory of a run (host code only).
//
// Layout and behaviour follow the reference's IO driver:
//  * field `id` is checkpointed to "<dir>/checkpoint/data<id>" when every
//    checkpoint of a run shares one file (params::single_output_file, the
//    default), else to "<dir>/checkpoint/data<id>_<index>"
//    (OUTPUT_CHECKPOINT_FMT / OUTPUT_CHECKPOINT_INDEX_FMT, io/inc/io.h:
//    137,185-188; io/src/writedefines.cpp:36-49);
//  * in the shared file the first checkpoint written by the process truncates
//    the file and carries the full GNU header; later ones are appended with a
//    header of the index alone, "<index> @" (io/src/writedefines.cpp:160-185,
//    io/src/writegp.cpp:427-465);
//  * each checkpoint appends "#$<index>" to "<dir>/checkpoint/progress.txt",
//    and "#$DONE" once the run reached its stop index (record_index,
//    modules-io.cpp:29-60; CONFIG_INDEX_PREFIX, conf/inc/confsystem.h:41,81);
//  * reading the shared file walks its blocks in order and keeps the last one
//    read, continuing while the next block's index does not pass the
//    requested one (read_grid_standardized, io/inc/readdefines.h:340-420).
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ckpt.hpp"

namespace sps {
namespace {

std::mutex g_mu;
// (dir, id) pairs whose shared checkpoint file this process has written
std::vector<std::pair<std::string, int>> g_written;

int make_dirs(const std::string& path) {
  std::string cur;
  size_t pos = 0;
  while (pos <= path.size()) {
    const size_t next = path.find('/', pos);
    cur = path.substr(0, next == std::string::npos ? path.size() : next);
    if (!cur.empty() && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) {
      set_error("cannot create directory '%s': %s", cur.c_str(),
                std::strerror(errno));
      return SPS_E_INVALID;
    }
    if (next == std::string::npos) break;
    pos = next + 1;
  }
  return SPS_OK;
}

std::string file_name(const char* dir, int id, int index, int single_file) {
  char buf[64];
  if (single_file)
    std::snprintf(buf, sizeof(buf), "/checkpoint/data%d", id);
  else
    std::snprintf(buf, sizeof(buf), "/checkpoint/data%d_%d", id, index);
  return std::string(dir) + buf;
}

int check_args(const char* dir, int id) {
  if (!dir || !*dir) {
    set_error("checkpoint directory is NULL or empty");
    return SPS_E_INVALID;
  }
  if (id < 0) {
    set_error("checkpoint field id %d is negative", id);
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

int read_file(const std::string& name, std::string& out) {
  FILE* f = std::fopen(name.c_str(), "rb");
  if (!f) {
    set_error("cannot open checkpoint file '%s': %s", name.c_str(),
              std::strerror(errno));
    return SPS_E_STATE;
  }
  out.clear();
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, got);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) {
    set_error("read error on checkpoint file '%s'", name.c_str());
    return SPS_E_STATE;
  }
  return SPS_OK;
}

// One line [pos, end of line) of s; returns the position after its newline
// (s.size() at the end).
size_t line_end(const std::string& s, size_t pos, size_t* content_end) {
  const size_t nl = s.find('\n', pos);
  if (nl == std::string::npos) {
    *content_end = s.size();
    return s.size();
  }
  *content_end = nl;
  return nl + 1;
}

// The full header "dim n0 .. n{dim-1} l0 r0 .. index tail": the rows per
// block (counts past the first axis multiplied) and the offset of the index
// token in the line.
bool parse_full_header(const std::string& line, long* rows, size_t* index_at) {
  const char* s = line.c_str();
  char* end = nullptr;
  const long dim = std::strtol(s, &end, 10);
  if (end == s || dim < 1 || dim > 3) return false;
  long r = 1;
  for (long a = 0; a < dim; ++a) {
    const char* p = end;
    const long n = std::strtol(p, &end, 10);
    if (end == p || n < 1) return false;
    if (a > 0) r *= n;
  }
  for (long a = 0; a < 2 * dim; ++a) {
    const char* p = end;
    std::strtod(p, &end);
    if (end == p) return false;
  }
  while (*end == ' ') ++end;
  *index_at = (size_t)(end - s);
  char* after = nullptr;
  std::strtol(end, &after, 10);
  if (after == end) return false;
  *rows = r;
  return true;
}

// A header line of a later block, "<index> @": its index, or -1 when the line
// is not one (the reference's BAD_INDEX, io/inc/readdefines.h:45).
int parse_short_header(const std::string& line) {
  const char* s = line.c_str();
  char* end = nullptr;
  const long v = std::strtol(s, &end, 10);
  if (end == s || v < 0 || v > 2147483647L) return -1;
  return (int)v;
}

}  // namespace

int ckpt_write_field(const char* dir, int id, int index, int single_file,
                     const CkptFormat& format) {
  SPS_TRY(check_args(dir, id));
  std::string text;
  SPS_TRY(format(index, text));
  SPS_TRY(make_dirs(std::string(dir) + "/checkpoint"));
  bool first = true;
  if (single_file) {
    std::lock_guard<std::mutex> lock(g_mu);
    const std::pair<std::string, int> key(dir, id);
    for (const auto& k : g_written)
      if (k == key) first = false;
    if (first) g_written.push_back(key);
  }
  const std::string name = file_name(dir, id, index, single_file);
  FILE* f = std::fopen(name.c_str(), first ? "wb" : "ab");
  if (!f) {
    set_error("cannot open checkpoint file '%s': %s", name.c_str(),
              std::strerror(errno));
    return SPS_E_INVALID;
  }
  size_t body = 0;
  if (!first) {
    // later blocks of the shared file: the header is the index alone
    const size_t nl = text.find('\n');
    body = nl == std::string::npos ? text.size() : nl + 1;
    std::fprintf(f, "%d @\n", index);
  }
  const size_t n = text.size() - body;
  const bool ok = std::fwrite(text.data() + body, 1, n, f) == n;
  if (std::fclose(f) != 0 || !ok) {
    set_error("write error on checkpoint file '%s'", name.c_str());
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

int ckpt_record(const char* dir, int index, int done) {
  SPS_TRY(check_args(dir, 0));
  SPS_TRY(make_dirs(std::string(dir) + "/checkpoint"));
  const std::string name = std::string(dir) + "/checkpoint/progress.txt";
  FILE* f = std::fopen(name.c_str(), "a");
  if (!f) {
    set_error("cannot open progress record '%s': %s", name.c_str(),
              std::strerror(errno));
    return SPS_E_INVALID;
  }
  std::fprintf(f, "#$%d\n", index);
  if (done) std::fprintf(f, "#$DONE\n");
  if (std::fclose(f) != 0) {
    set_error("write error on progress record '%s'", name.c_str());
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

int ckpt_read_field(const char* dir, int id, int index, int single_file,
                    int* index_read, const CkptParse& parse) {
  SPS_TRY(check_args(dir, id));
  std::string s;
  SPS_TRY(read_file(file_name(dir, id, index, single_file), s));
  if (!single_file) return parse(s.data(), s.size(), index_read);
  size_t ce = 0;
  size_t pos = line_end(s, 0, &ce);
  const std::string head = s.substr(0, ce);
  long rows = 0;
  size_t index_at = 0;
  if (!parse_full_header(head, &rows, &index_at)) {
    set_error("checkpoint file of field %d: malformed header", id);
    return SPS_E_INVALID;
  }
  int cur = std::atoi(head.c_str() + index_at);
  // the reference's loop: read the block, then the next header; go on while
  // that index does not pass the requested one and the last block read was
  // not the requested one
  size_t sel_begin = 0, sel_end = 0;
  int prev = -1;
  for (;;) {
    const size_t begin = pos;
    for (long r = 0; r < rows && pos < s.size(); ++r) pos = line_end(s, pos, &ce);
    sel_begin = begin;
    sel_end = pos;
    prev = cur;
    int next = -1;
    if (pos < s.size()) {
      const size_t after = line_end(s, pos, &ce);
      next = parse_short_header(s.substr(pos, ce - pos));
      pos = after;
    }
    cur = next;
    if (!(cur <= index && cur > -1 && prev != index)) break;
  }
  // the selected block under the file's full header, carrying its own index
  std::string text = head.substr(0, index_at) + std::to_string(prev) + " @\n";
  text.append(s, sel_begin, sel_end - sel_begin);
  return parse(text.data(), text.size(), index_read);
}

}  // namespace sps

using namespace sps;

extern "C" int sps_fd_checkpoint_save(sps_fd_plan* plan, const char* dir,
                                      int id, int index, int single_file,
                                      int done, const double* origin) {
  if (!plan) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  SPS_TRY(ckpt_write_field(
      dir, id, index, single_file, [&](int idx, std::string& text) {
        size_t len = 0;
        SPS_TRY(sps_fd_checkpoint_text(plan, idx, origin, nullptr, 0, &len));
        text.assign(len, '\0');
        return sps_fd_checkpoint_text(plan, idx, origin, &text[0], len, &len);
      }));
  return ckpt_record(dir, index, done);
}

extern "C" int sps_fd_checkpoint_load(sps_fd_plan* plan, const char* dir,
                                      int id, int index, int single_file,
                                      int* index_read) {
  if (!plan) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  return ckpt_read_field(dir, id, index, single_file, index_read,
                         [&](const char* text, size_t len, int* idx) {
                           return sps_fd_checkpoint_read(plan, text, len, idx);
                         });
}

extern "C" int sps_fd_checkpoint_progress(const char* dir, int* last_index,
                                          int* count, int* done) {
  SPS_TRY(check_args(dir, 0));
  if (last_index) *last_index = -1;
  if (count) *count = 0;
  if (done) *done = 0;
  const std::string name = std::string(dir) + "/checkpoint/progress.txt";
  FILE* f = std::fopen(name.c_str(), "r");
  if (!f) return SPS_OK;  // no checkpoint recorded yet
  char line[256];
  while (std::fgets(line, sizeof(line), f)) {
    // trim the line as str_trim does, then take what follows "#$"
    char* b = line;
    while (*b == ' ' || *b == '\t') ++b;
    char* e = b + std::strlen(b);
    while (e > b && (e[-1] == '\n' || e[-1] == '\r' || e[-1] == ' ' ||
                     e[-1] == '\t'))
      *--e = '\0';
    if (std::strncmp(b, "#$", 2) != 0) continue;
    b += 2;
    char* end = nullptr;
    const long v = std::strtol(b, &end, 10);
    if (end != b && !*end) {
      if (last_index) *last_index = (int)v;
      if (count) ++*count;
    } else if (std::strcmp(b, "DONE") == 0) {
      if (done) *done = 1;
    }
  }
  std::fclose(f);
  return SPS_OK;
}
