// Host I/O on either side of the force path, in the reference's on-disk
// formats: the particle CSV read by read_csv (ttg-parsec/src/io.h:10-14) and
// the legacy-VTK snapshot written by write_vtk (io.h:16-18).
//
// CSV records are "id,mass,x,y,z,vx,vy,vz", one per line. Blanks and tabs may
// surround a field, a trailing '\r' is ignored, lines holding only blanks are
// skipped, and a first line whose first field is not an unsigned integer is a
// header. Any other line that is not a record is an error; the message gives
// its line number and text.
//
// Reader design: the file is read into memory and cut into fixed 8 MiB
// blocks. A line belongs to the block it starts in. Worker threads claim
// blocks from an atomic counter; each splits its lines into eight field
// spans with one scan for separators, then converts the spans. The blocks'
// records are then copied, in block order, into one array.
#include <fcntl.h>
#include <sched.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
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

constexpr size_t kBlockBytes = size_t(8) << 20;
constexpr int kFields = 8;

void put_err(char* buf, size_t len, const std::string& msg) {
  if (buf && len) std::snprintf(buf, len, "%s", msg.c_str());
}

bool is_blank(char c) { return c == ' ' || c == '\t'; }

// A field's text without surrounding blanks.
struct Span {
  const char* b;
  const char* e;
};

Span trim(const char* b, const char* e) {
  while (b < e && is_blank(*b)) ++b;
  while (e > b && is_blank(e[-1])) --e;
  return {b, e};
}

// Splits [b, e) (one line, no '\n') at commas. Returns the number of fields
// found (kFields + 1 means "too many").
int split(const char* b, const char* e, Span* f) {
  if (e > b && e[-1] == '\r') --e;
  int k = 0;
  const char* start = b;
  for (const char* p = b;; ++p) {
    if (p == e || *p == ',') {
      if (k == kFields) return kFields + 1;
      f[k++] = trim(start, p);
      if (p == e) break;
      start = p + 1;
    }
  }
  return k;
}

template <typename T>
bool convert(Span s, T& v) {
  if (s.b == s.e) return false;
  const auto r = std::from_chars(s.b, s.e, v);
  return r.ec == std::errc() && r.ptr == s.e;
}

bool blank_line(const char* b, const char* e) {
  for (; b < e; ++b)
    if (!is_blank(*b) && *b != '\r') return false;
  return true;
}

// One parsed block: its records and, if any, its first bad line.
struct Block {
  std::vector<bh_particle> recs;
  const char* bad = nullptr;
  const char* bad_end = nullptr;
};

// Parses the lines that start in [lo, hi) (the last one may run past hi,
// up to end).
void parse_block(const char* lo, const char* hi, const char* end, Block& out) {
  Span f[kFields];
  const char* line = lo;
  while (line < hi) {
    const char* nl = static_cast<const char*>(std::memchr(line, '\n', static_cast<size_t>(end - line)));
    const char* eol = nl ? nl : end;
    if (!blank_line(line, eol)) {
      bh_particle r{};
      const bool ok = split(line, eol, f) == kFields && convert(f[0], r.id) && convert(f[1], r.mass) &&
                      convert(f[2], r.pos[0]) && convert(f[3], r.pos[1]) && convert(f[4], r.pos[2]) &&
                      convert(f[5], r.vel[0]) && convert(f[6], r.vel[1]) && convert(f[7], r.vel[2]);
      if (!ok) {
        out.bad = line;
        out.bad_end = eol > line && eol[-1] == '\r' ? eol - 1 : eol;
        return;
      }
      out.recs.push_back(r);
    }
    line = eol + 1;
  }
}

// First byte of the first line starting at or after p (a line starts at the
// data's beginning or right after a '\n').
const char* line_start_at_or_after(const char* data, const char* p, const char* end) {
  if (p <= data) return data;
  if (p >= end) return end;
  if (p[-1] == '\n') return p;
  const char* nl = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
  return nl ? nl + 1 : end;
}

unsigned worker_count(size_t blocks) {
  unsigned cpus = std::thread::hardware_concurrency();
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = static_cast<unsigned>(CPU_COUNT(&set));
  return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(blocks, std::max(1u, cpus))));
}

bool slurp(const char* path, std::vector<char>& buf, std::string& why) {
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) {
    why = std::string("cannot open input file: ") + path;
    return false;
  }
  struct stat st {};
  if (::fstat(fd, &st) != 0) {
    ::close(fd);
    why = std::string("cannot stat input file: ") + path;
    return false;
  }
  buf.resize(static_cast<size_t>(st.st_size));
  size_t got = 0;
  while (got < buf.size()) {
    const ssize_t r = ::read(fd, buf.data() + got, buf.size() - got);
    if (r <= 0) {
      ::close(fd);
      why = std::string("read failed: ") + path;
      return false;
    }
    got += static_cast<size_t>(r);
  }
  ::close(fd);
  return true;
}

}  // namespace

extern "C" {

int bh_read_csv(const char* path, bh_particle** out, uint64_t* n, char* err, size_t err_len) {
  if (!path || !out || !n) return BH_ERR_INVALID;
  *out = nullptr;
  *n = 0;
  std::vector<char> file;
  std::string why;
  if (!slurp(path, file, why)) {
    put_err(err, err_len, why);
    return BH_ERR_INVALID;
  }
  const char* data = file.data();
  const char* end = data + file.size();
  const char* body = data;
  if (!file.empty()) {
    // header: the first line's first field is not an unsigned integer
    const char* nl = static_cast<const char*>(std::memchr(data, '\n', file.size()));
    const char* eol = nl ? nl : end;
    const char* comma = static_cast<const char*>(std::memchr(data, ',', static_cast<size_t>(eol - data)));
    uint32_t id = 0;
    if (!convert(trim(data, comma ? comma : eol), id)) body = nl ? nl + 1 : end;
  }
  const size_t bytes = static_cast<size_t>(end - body);
  const size_t blocks = bytes == 0 ? 0 : (bytes + kBlockBytes - 1) / kBlockBytes;
  std::vector<Block> part(blocks);
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t k; (k = next.fetch_add(1, std::memory_order_relaxed)) < blocks;) {
      const char* lo = line_start_at_or_after(data, body + k * kBlockBytes, end);
      const char* hi = line_start_at_or_after(data, body + std::min(bytes, (k + 1) * kBlockBytes), end);
      parse_block(lo, hi, end, part[k]);
    }
  };
  {
    std::vector<std::thread> pool(worker_count(blocks));
    for (auto& t : pool) t = std::thread(work);
    for (auto& t : pool) t.join();
  }
  for (const Block& b : part) {
    if (b.bad) {
      const long line_no = 1 + std::count(data, b.bad, '\n');
      put_err(err, err_len,
              std::string("malformed line ") + std::to_string(line_no) + " of " + path + ": '" +
                  std::string(b.bad, b.bad_end) + "'");
      return BH_ERR_INVALID;
    }
  }
  uint64_t total = 0;
  for (const Block& b : part) total += b.recs.size();
  if (total == 0) return BH_OK;
  auto* dst = static_cast<bh_particle*>(std::malloc(total * sizeof(bh_particle)));
  if (!dst) {
    put_err(err, err_len, "out of host memory");
    return BH_ERR_OOM;
  }
  bh_particle* w = dst;
  for (const Block& b : part) {
    if (!b.recs.empty()) std::memcpy(w, b.recs.data(), b.recs.size() * sizeof(bh_particle));
    w += b.recs.size();
  }
  *out = dst;
  *n = total;
  return BH_OK;
}

void bh_free(void* p) { std::free(p); }

// Legacy VTK ASCII PolyData: the points, then per-point data with a mass
// scalar and a velocity vector, doubles in the stream's default format.
int bh_write_vtk(const char* path, const bh_particle* p, uint64_t n, char* err, size_t err_len) {
  if (!path || (n && !p)) return BH_ERR_INVALID;
  std::ofstream os(path);
  if (!os) {
    put_err(err, err_len, std::string("cannot open output file: ") + path);
    return BH_ERR_INVALID;
  }
  auto triples = [&](auto get) {
    for (uint64_t i = 0; i < n; ++i) {
      const double* v = get(p[i]);
      os << v[0] << ' ' << v[1] << ' ' << v[2] << '\n';
    }
  };
  os << "# vtk DataFile Version 3.0\n"
     << "pepc-ttg particle snapshot\n"
     << "ASCII\n"
     << "DATASET POLYDATA\n"
     << "POINTS " << n << " double\n";
  triples([](const bh_particle& q) { return q.pos; });
  os << "POINT_DATA " << n << '\n' << "SCALARS mass double 1\n" << "LOOKUP_TABLE default\n";
  for (uint64_t i = 0; i < n; ++i) os << p[i].mass << '\n';
  os << "VECTORS velocity double\n";
  triples([](const bh_particle& q) { return q.vel; });
  if (!os.flush()) {
    put_err(err, err_len, std::string("write failed: ") + path);
    return BH_ERR_INVALID;
  }
  return BH_OK;
}

}  // extern "C"
