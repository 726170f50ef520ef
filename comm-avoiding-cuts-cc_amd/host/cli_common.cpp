// cli_common.cpp -- see cli_common.h.
#include "cli_common.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "cac_common_host.h"

namespace cli {
namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Parses one unsigned decimal starting at p (after skipping whitespace).
inline bool next_u64(const char*& p, const char* end, uint64_t* v) {
  while (p < end && is_space(*p)) ++p;
  if (p >= end || *p < '0' || *p > '9') return false;
  uint64_t x = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    x = x * 10 + (uint64_t)(*p - '0');
    ++p;
  }
  *v = x;
  return true;
}

// Optional binary cache of a parsed edge list, beside the text file as
// <path>.cacbin: this 64-byte header, then m cac_edge records.  It is used only
// while the text file keeps the size and modification time recorded here.
struct CacheHeader {
  char magic[8];  // "CACEDGE1"
  uint32_t n, pad;
  uint64_t m, text_size;
  int64_t text_mtime_ns;
  uint64_t reserved[3];
};
static_assert(sizeof(CacheHeader) == 64, "cache header layout");
constexpr char kCacheMagic[8] = {'C', 'A', 'C', 'E', 'D', 'G', 'E', '1'};

bool cache_enabled() {
  const char* e = std::getenv("CAC_EDGE_CACHE");
  return e && *e && std::strcmp(e, "0") != 0;
}

int64_t mtime_ns(const struct stat& st) { return (int64_t)st.st_mtim.tv_sec * 1000000000 + st.st_mtim.tv_nsec; }

bool read_all(int fd, void* buf, size_t bytes) {
  char* p = (char*)buf;
  while (bytes) {
    const ssize_t r = ::read(fd, p, bytes);
    if (r <= 0) return false;
    p += r;
    bytes -= (size_t)r;
  }
  return true;
}

bool write_all(int fd, const void* buf, size_t bytes) {
  const char* p = (const char*)buf;
  while (bytes) {
    const ssize_t r = ::write(fd, p, bytes);
    if (r <= 0) return false;
    p += r;
    bytes -= (size_t)r;
  }
  return true;
}

bool read_cache(const std::string& path, const struct stat& text, uint32_t* n, std::vector<cac_edge>* edges) {
  const int fd = ::open((path + ".cacbin").c_str(), O_RDONLY);
  if (fd < 0) return false;
  CacheHeader h;
  struct stat st;
  bool ok = fstat(fd, &st) == 0 && read_all(fd, &h, sizeof h) && std::memcmp(h.magic, kCacheMagic, 8) == 0 &&
            h.text_size == (uint64_t)text.st_size && h.text_mtime_ns == mtime_ns(text) &&
            (uint64_t)st.st_size == sizeof h + h.m * sizeof(cac_edge);
  if (ok) {
    edges->resize(h.m);
    ok = read_all(fd, edges->data(), h.m * sizeof(cac_edge));
    if (ok) *n = h.n;
  }
  ::close(fd);
  return ok;
}

// Best effort: written to a private temporary name, then renamed into place.
void write_cache(const std::string& path, const struct stat& text, uint32_t n, const std::vector<cac_edge>& edges) {
  CacheHeader h{};
  std::memcpy(h.magic, kCacheMagic, 8);
  h.n = n;
  h.m = edges.size();
  h.text_size = (uint64_t)text.st_size;
  h.text_mtime_ns = mtime_ns(text);
  const std::string tmp = path + ".cacbin.tmp" + std::to_string((long)getpid());
  const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) return;
  const bool ok = write_all(fd, &h, sizeof h) && write_all(fd, edges.data(), edges.size() * sizeof(cac_edge));
  ::close(fd);
  if (!ok || std::rename(tmp.c_str(), (path + ".cacbin").c_str()) != 0) ::unlink(tmp.c_str());
}

}  // namespace

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool parse_float(const char* s, float* out) {
  try {
    *out = std::stof(s, nullptr);
    return true;
  } catch (...) {
    return false;
  }
}

bool parse_int(const char* s, int* out) {
  try {
    *out = std::stoi(s);
    return true;
  } catch (...) {
    return false;
  }
}

int read_edge_list(const std::string& path, uint32_t* n_out, std::vector<cac_edge>* edges, std::string* err) {
  int fd = ::open(path.c_str(), O_RDONLY);
  if (fd < 0) {
    *err = "cannot open " + path;
    return CAC_EIO;
  }
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size == 0) {
    ::close(fd);
    *err = "cannot read " + path;
    return CAC_EIO;
  }
  const bool cache = cache_enabled();
  if (cache && read_cache(path, st, n_out, edges)) {
    ::close(fd);
    return CAC_OK;
  }
  const size_t size = (size_t)st.st_size;
  const char* base = (const char*)mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (base == MAP_FAILED) {
    *err = "cannot map " + path;
    return CAC_EIO;
  }
  const char* end = base + size;
  const char* p = (const char*)memchr(base, '\n', size);
  int rc = CAC_OK;
  uint64_t n = 0, m = 0;
  if (!p || !next_u64(++p, end, &n) || !next_u64(p, end, &m) || n > 0xFFFFFFFFull) {
    *err = "malformed header in " + path;
    rc = CAC_EIO;
  }
  if (rc == CAC_OK) {
    // split the body into chunks at line boundaries; each chunk is parsed by one
    // thread into its own vector, then concatenated in order
    while (p < end && *p != '\n') ++p;  // rest of the "n m" line
    const unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t body = (size_t)(end - p);
    std::vector<const char*> cut(T + 1);
    cut[0] = p;
    cut[T] = end;
    for (unsigned t = 1; t < T; ++t) {
      const char* q = p + body * t / T;
      if (q < cut[t - 1]) q = cut[t - 1];
      while (q < end && *q != '\n') ++q;
      cut[t] = q;
    }
    std::vector<std::vector<cac_edge>> part(T);
    std::vector<int> ok(T, 1);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t) {
      th.emplace_back([&, t]() {
        const char* q = cut[t];
        const char* e = cut[t + 1];
        std::vector<cac_edge>& out = part[t];
        out.reserve((size_t)(m / T + 16));
        uint64_t a, b, w;
        for (;;) {
          const char* save = q;
          if (!next_u64(q, e, &a)) {
            while (save < e && is_space(*save)) ++save;
            if (save != e) ok[t] = 0;  // trailing garbage
            break;
          }
          if (!next_u64(q, e, &b) || !next_u64(q, e, &w) || a > 0xFFFFFFFFull || b > 0xFFFFFFFFull) {
            ok[t] = 0;
            break;
          }
          out.push_back(cac_edge{(uint32_t)a, (uint32_t)b, w});
        }
      });
    }
    for (auto& x : th) x.join();
    size_t total = 0;
    for (unsigned t = 0; t < T; ++t) {
      if (!ok[t]) rc = CAC_EIO;
      total += part[t].size();
    }
    if (rc != CAC_OK) *err = "malformed edge line in " + path;
    else if (total < m) {
      *err = "fewer edge lines than announced in " + path;
      rc = CAC_EIO;
    } else {
      edges->clear();
      edges->reserve(m);
      for (unsigned t = 0; t < T && edges->size() < m; ++t)
        for (const cac_edge& x : part[t]) {
          if (edges->size() == m) break;
          edges->push_back(x);
        }
      *n_out = (uint32_t)n;
      if (cache) write_cache(path, st, *n_out, *edges);
    }
  }
  munmap((void*)base, size);
  return rc;
}

std::vector<cac_edge> click_graph(uint32_t n, uint32_t seed) {
  std::vector<cac_edge> e;
  e.reserve((size_t)n * (n - 1) / 2);
  uint64_t idx = 0;
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j, ++idx) {
      const cac::u32x4 r = cac::philox4x32((uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u, seed, cac::kDomGen);
      // Box-Muller on two 32-bit uniforms in (0, 1]
      const double u1 = ((double)r.x + 1.0) / 4294967296.0, u2 = ((double)r.y + 1.0) / 4294967296.0;
      const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      const double mu = (i % 10 == j % 10) ? 8.0 : 4.0;
      const double w = std::max(0.0, mu + 4.0 * z);
      e.push_back(cac_edge{i, j, (uint64_t)w});
    }
  return e;
}

}  // namespace cli
