// Host I/O on either side of the force path: the particle CSV input
// (ttg-parsec/src/io.h:10-14, io.cc:94-178) and the legacy-VTK snapshot
// (io.h:16-18, io.cc:180-195). The formats are the reference's; the code is
// this library's own.
//
// CSV: one particle per line, "id,mass,x,y,z,vx,vy,vz", blanks around fields
// allowed, CRLF accepted, blank lines skipped; a first line whose first
// field is not an integer is a header and skipped; any other malformed line
// is an error naming that line. Large files are parsed in parallel chunks
// split at line starts.
#include <fcntl.h>
#include <sched.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bh_gpu.h"

namespace {

void put_err(char* buf, size_t len, const std::string& msg) {
  if (buf && len) {
    std::snprintf(buf, len, "%s", msg.c_str());
  }
}

unsigned cpus() {
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
  return std::max(1u, std::thread::hardware_concurrency());
}

inline const char* blanks(const char* p, const char* e) {
  while (p < e && (*p == ' ' || *p == '\t')) ++p;
  return p;
}

template <typename T>
bool field(const char*& p, const char* e, T& v, bool last) {
  p = blanks(p, e);
  const auto r = std::from_chars(p, e, v);
  if (r.ec != std::errc()) return false;
  const char* q = blanks(r.ptr, e);
  if (last) {
    if (q < e && *q == '\r') ++q;
    if (q != e) return false;
  } else {
    if (q == e || *q != ',') return false;
    ++q;
  }
  p = q;
  return true;
}

// Parses whole lines in [b, e); returns the first malformed line or "".
std::string parse(const char* b, const char* e, std::vector<bh_particle>& out) {
  for (const char* line = b; line < e;) {
    const char* eol = static_cast<const char*>(std::memchr(line, '\n', static_cast<size_t>(e - line)));
    if (!eol) eol = e;
    const char* p = blanks(line, eol);
    if (p != eol && *p != '\r') {
      bh_particle q{};
      if (!(field(p, eol, q.id, false) && field(p, eol, q.mass, false) && field(p, eol, q.pos[0], false) &&
            field(p, eol, q.pos[1], false) && field(p, eol, q.pos[2], false) && field(p, eol, q.vel[0], false) &&
            field(p, eol, q.vel[1], false) && field(p, eol, q.vel[2], true)))
        return std::string(line, eol);
      out.push_back(q);
    }
    line = eol + 1;
  }
  return {};
}

}  // namespace

extern "C" {

int bh_read_csv(const char* path, bh_particle** out, uint64_t* n, char* err, size_t err_len) {
  if (!path || !out || !n) return BH_ERR_INVALID;
  *out = nullptr;
  *n = 0;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) {
    put_err(err, err_len, std::string("cannot open input file: ") + path);
    return BH_ERR_INVALID;
  }
  struct stat st {};
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    put_err(err, err_len, std::string("cannot stat input file: ") + path);
    return BH_ERR_INVALID;
  }
  const size_t size = static_cast<size_t>(st.st_size);
  const char* data = nullptr;
  if (size > 0) {
    void* m = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
      ::close(fd);
      put_err(err, err_len, std::string("cannot map input file: ") + path);
      return BH_ERR_INVALID;
    }
    data = static_cast<const char*>(m);
  }
  ::close(fd);
  const char* b = data;
  const char* e = data + size;
  if (b != e) {
    const char* eol = static_cast<const char*>(std::memchr(b, '\n', size));
    const char* p = blanks(b, eol ? eol : e);
    uint32_t id;
    if (std::from_chars(p, eol ? eol : e, id).ec != std::errc()) b = eol ? eol + 1 : e;  // header
  }
  const size_t bytes = static_cast<size_t>(e - b);
  const unsigned chunks = static_cast<unsigned>(std::clamp<size_t>(bytes >> 25, 1, cpus()));
  std::vector<const char*> cut(chunks + 1);
  cut[0] = b;
  cut[chunks] = e;
  for (unsigned c = 1; c < chunks; ++c) {
    const char* p = std::max(b + bytes * c / chunks, cut[c - 1]);
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(e - p)));
    cut[c] = eol ? eol + 1 : e;
  }
  std::vector<std::vector<bh_particle>> part(chunks);
  std::vector<std::string> bad(chunks);
  {
    std::vector<std::thread> th;
    for (unsigned c = 1; c < chunks; ++c) th.emplace_back([&, c] { bad[c] = parse(cut[c], cut[c + 1], part[c]); });
    bad[0] = parse(cut[0], cut[1], part[0]);
    for (auto& t : th) t.join();
  }
  if (data) ::munmap(const_cast<char*>(data), size);
  for (const auto& s : bad) {
    if (!s.empty()) {
      put_err(err, err_len, std::string("malformed line in ") + path + ": '" + s + "'");
      return BH_ERR_INVALID;
    }
  }
  uint64_t total = 0;
  for (const auto& v : part) total += v.size();
  if (total == 0) return BH_OK;
  bh_particle* buf = static_cast<bh_particle*>(std::malloc(total * sizeof(bh_particle)));
  if (!buf) {
    put_err(err, err_len, "out of host memory");
    return BH_ERR_OOM;
  }
  uint64_t off = 0;
  for (const auto& v : part) {
    std::memcpy(buf + off, v.data(), v.size() * sizeof(bh_particle));
    off += v.size();
  }
  *out = buf;
  *n = total;
  return BH_OK;
}

void bh_free(void* p) { std::free(p); }

int bh_write_vtk(const char* path, const bh_particle* p, uint64_t n, char* err, size_t err_len) {
  if (!path || (n && !p)) return BH_ERR_INVALID;
  std::ofstream out(path);
  if (!out) {
    put_err(err, err_len, std::string("cannot open output file: ") + path);
    return BH_ERR_INVALID;
  }
  out << "# vtk DataFile Version 3.0\npepc-ttg particle snapshot\nASCII\nDATASET POLYDATA\n";
  out << "POINTS " << n << " double\n";
  for (uint64_t i = 0; i < n; ++i) out << p[i].pos[0] << ' ' << p[i].pos[1] << ' ' << p[i].pos[2] << '\n';
  out << "POINT_DATA " << n << "\nSCALARS mass double 1\nLOOKUP_TABLE default\n";
  for (uint64_t i = 0; i < n; ++i) out << p[i].mass << '\n';
  out << "VECTORS velocity double\n";
  for (uint64_t i = 0; i < n; ++i) out << p[i].vel[0] << ' ' << p[i].vel[1] << ' ' << p[i].vel[2] << '\n';
  if (!out) {
    put_err(err, err_len, std::string("write failed: ") + path);
    return BH_ERR_INVALID;
  }
  return BH_OK;
}

}  // extern "C"
