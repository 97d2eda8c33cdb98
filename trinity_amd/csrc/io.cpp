// io.cpp -- mesh and field input, replacing the parse of Mesh::load
// (src/io.cpp:25-48, :53-114) and adding a binary format.
//
// Text (.mesh / .bb, the reference's input): the file is mapped and each
// section is parsed by several threads.  A first pass over byte chunks
// counts the non-empty lines that start in each chunk; their prefix sums
// give every line its record number, and a second pass parses the lines of
// each chunk into their records.  The semantics are Mesh::load's: sections
// are searched in the order Vertices, Edges, Triangles, Corners, each from
// where the previous one ended, a keyword being the first token of a line
// that does not start with '#' or '%' (io::find), and the count the next
// token; vertices are "x y ref", edges "a b ref" (both ends tagged
// boundary), triangles "a b c ref" (1-based, stored 0-based), corners "a"
// (tagged boundary | corner, after the edges); the .bb values start on its
// second line (tools::seekToLine(2)).  Numbers are read with correctly
// rounded conversions, so coordinates and values equal the reference's
// bit for bit.  Records must be one per line (every file the reference
// writes and ships is); a line holding more or fewer numbers than its
// record is an error, as is an index outside 1..nv, rather than a silent
// misread.
//
// Binary (TRNM, this library's own): "TRNM", int32 version 1, int32 nv,
// int32 nt, then f64 xy[2nv], int32 tri[3nt] (0-based), u8 tags[nv], f64
// solut[nv], little-endian.  Reads go straight into host arrays, or into
// device arrays through two pinned staging buffers whose copies overlap the
// next read.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "trn_api.hpp"

namespace trn {
namespace {

constexpr uint8_t kTagBound = 4;    // mask::bound (include/trinity/header.h:57)
constexpr uint8_t kTagCorner = 8;   // mask::corner

int io_threads() {
  if (const char* e = std::getenv("TRN_IO_THREADS")) {
    const int n = std::atoi(e);
    if (n > 0) return std::min(n, 256);
  }
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
}

// Read-only mapping of a whole file.
class MappedFile {
 public:
  explicit MappedFile(const char* path) {
    TRN_REQUIRE(path != nullptr, TRN_ERR_ARG, "null path");
    fd_ = ::open(path, O_RDONLY);
    TRN_REQUIRE(fd_ >= 0, TRN_ERR_ARG, std::string("cannot open ") + path);
    struct stat st;
    if (::fstat(fd_, &st) != 0) {
      ::close(fd_);
      fd_ = -1;
      TRN_REQUIRE(false, TRN_ERR_ARG, std::string("cannot stat ") + path);
    }
    size_ = (size_t)st.st_size;
    if (size_ > 0) {
      void* p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p == MAP_FAILED) {
        ::close(fd_);
        fd_ = -1;
        TRN_REQUIRE(false, TRN_ERR_ARG, std::string("cannot map ") + path);
      }
      (void)::madvise(p, size_, MADV_SEQUENTIAL);
      data_ = static_cast<const char*>(p);
    }
  }
  MappedFile(const MappedFile&) = delete;
  MappedFile& operator=(const MappedFile&) = delete;
  ~MappedFile() {
    if (data_) ::munmap(const_cast<char*>(data_), size_);
    if (fd_ >= 0) ::close(fd_);
  }
  const char* data() const { return data_; }
  size_t size() const { return size_; }

 private:
  int fd_ = -1;
  const char* data_ = nullptr;
  size_t size_ = 0;
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

// Start of the line after position p (end if none).
size_t next_line(const char* d, size_t n, size_t p) {
  const void* q = p < n ? std::memchr(d + p, '\n', n - p) : nullptr;
  return q ? (size_t)((const char*)q - d) + 1 : n;
}

// io::find: from the line at p on, the first line (comments skipped) whose
// first token is key; returns the position after that line, or n.
size_t find_key(const char* d, size_t n, size_t p, const char* key) {
  const size_t kl = std::strlen(key);
  while (p < n) {
    const size_t e = next_line(d, n, p);
    size_t s = p;
    while (s < e && is_space(d[s])) ++s;
    if (s < e && d[s] != '#' && d[s] != '%') {
      size_t t = s;
      while (t < e && !is_space(d[t])) ++t;
      if (t - s == kl && std::memcmp(d + s, key, kl) == 0) return e;
    }
    p = e;
  }
  return n;
}

// Next token at or after p as an integer (the section count); p moves past it.
bool read_count(const char* d, size_t n, size_t& p, int64_t& out) {
  while (p < n && is_space(d[p])) ++p;
  if (p >= n) return false;
  const auto r = std::from_chars(d + p, d + n, out);
  if (r.ec != std::errc()) return false;
  p = (size_t)(r.ptr - d);
  return true;
}

// A record's numbers: up to 4 per line, parsed as doubles or integers.
struct Fields {
  double f[4];
  int64_t i[4];
};

// Parses exactly `want` numbers from [s, e) (the first `nf` as doubles, the
// rest as integers); false if the line holds another count or a bad token.
bool parse_line(const char* s, const char* e, int want, int nf, Fields& out) {
  int k = 0;
  while (true) {
    while (s < e && is_space(*s)) ++s;
    if (s >= e) break;
    if (k == want) return false;
    std::from_chars_result r;
    if (k < nf) r = std::from_chars(s, e, out.f[k]);
    else r = std::from_chars(s, e, out.i[k]);
    if (r.ec != std::errc() || (r.ptr < e && !is_space(*r.ptr))) return false;
    s = r.ptr;
    ++k;
  }
  return k == want;
}

// Start of the records of a section whose count token ended at p: the rest
// of the count's line must be blank, records start on the next line.
size_t records_start(const char* d, size_t n, size_t p, const char* what) {
  const size_t e = next_line(d, n, p);
  for (size_t q = p; q < e; ++q)
    TRN_REQUIRE(is_space(d[q]), TRN_ERR_ARG,
                std::string(what) + ": records must start on the line after the count");
  return e;
}

// One section: `count` records, one per non-empty line, the first at or
// after line start r0.  store(record, fields) is called for every record
// (from several threads; records are disjoint) and returns false for a
// bad one.  Lines are counted in rounds of io_threads() chunks of
// kChunkBytes until the section's lines are all found, then the chunks
// holding them are parsed.  Returns the position after the last record's
// line.
constexpr size_t kChunkBytes = size_t(4) << 20;

template <class Store>
size_t parse_section(const char* d, size_t n, size_t r0, int64_t count, int want, int nf,
                     const char* what, Store store) {
  if (count <= 0) return r0;
  const int T = io_threads();
  std::vector<size_t> cs{r0};     // chunk starts (line starts), then the end
  std::vector<int64_t> lines{0};  // non-empty lines before each chunk
  auto run = [&](int lo, int hi, auto&& fn) {
    std::atomic<int> next{lo};
    std::vector<std::thread> th;
    for (int t = 0; t < std::min(T, hi - lo); ++t)
      th.emplace_back([&] {
        for (int c = next.fetch_add(1); c < hi; c = next.fetch_add(1)) fn(c);
      });
    for (auto& t : th) t.join();
  };
  while (lines.back() < count && cs.back() < n) {
    const int g0 = (int)cs.size() - 1;
    for (int t = 0; t < T && cs.back() < n; ++t)
      cs.push_back(next_line(d, n, std::min(n, cs.back() + kChunkBytes) - 1));
    const int g1 = (int)cs.size() - 1;
    std::vector<int64_t> cnt(g1 - g0, 0);
    run(g0, g1, [&](int c) {
      int64_t k = 0;
      for (size_t q = cs[c]; q < cs[c + 1];) {
        const size_t e = next_line(d, n, q);
        size_t s = q;
        while (s < e && is_space(d[s])) ++s;
        if (s < e) ++k;
        q = e;
      }
      cnt[c - g0] = k;
    });
    for (int c = g0; c < g1; ++c) lines.push_back(lines.back() + cnt[c - g0]);
  }
  TRN_REQUIRE(lines.back() >= count, TRN_ERR_ARG,
              std::string(what) + ": file ends after " + std::to_string(lines.back()) +
                  " of " + std::to_string(count) + " records");
  const int nc = (int)cs.size() - 1;
  std::atomic<int64_t> bad{0};
  std::atomic<size_t> end_pos{n};
  run(0, nc, [&](int c) {
    int64_t rec = lines[c];
    if (rec >= count) return;
    Fields fl;
    for (size_t q = cs[c]; q < cs[c + 1] && rec < count;) {
      const size_t e = next_line(d, n, q);
      size_t s = q;
      while (s < e && is_space(d[s])) ++s;
      if (s < e) {
        if (!parse_line(d + s, d + e, want, nf, fl) || !store(rec, fl)) bad.fetch_add(1);
        if (++rec == count) end_pos.store(e);
      }
      q = e;
    }
  });
  TRN_REQUIRE(bad.load() == 0, TRN_ERR_ARG,
              std::string(what) + ": " + std::to_string(bad.load()) +
                  " records are not one per line with the expected numbers, or hold an "
                  "index out of range");
  return end_pos.load();
}

struct MeditCounts {
  int64_t nv = -1, nt = -1;
};

MeditCounts medit_counts(const MappedFile& f) {
  MeditCounts c;
  const char* d = f.data();
  const size_t n = f.size();
  size_t p = find_key(d, n, 0, "Vertices");
  if (p < n && read_count(d, n, p, c.nv)) {
    // Triangles are searched after Edges, as Mesh::load does
    size_t q = find_key(d, n, p, "Edges");
    q = find_key(d, n, q, "Triangles");
    if (q < n) read_count(d, n, q, c.nt);
  }
  return c;
}

}  // namespace
}  // namespace trn

using namespace trn;

extern "C" {

int trn_medit_info(const char* mesh_path, int32_t* nv, int32_t* nt) {
  return guard_impl("trn_medit_info", [&] {
    TRN_REQUIRE(nv && nt, TRN_ERR_ARG, "null output");
    MappedFile f(mesh_path);
    const MeditCounts c = medit_counts(f);
    TRN_REQUIRE(c.nv > 0 && c.nv < INT32_MAX, TRN_ERR_ARG, "no Vertices section");
    TRN_REQUIRE(c.nt > 0 && c.nt < INT32_MAX / 3, TRN_ERR_ARG, "no Triangles section");
    *nv = (int32_t)c.nv;
    *nt = (int32_t)c.nt;
  });
}

int trn_medit_read(const char* mesh_path, const char* bb_path, int32_t nv, int32_t nt,
                   double* xy, int32_t* tri, uint8_t* tags, double* solut) {
  return guard_impl("trn_medit_read", [&] {
    TRN_REQUIRE(xy && tri && tags, TRN_ERR_ARG, "null output array");
    TRN_REQUIRE(!solut || bb_path, TRN_ERR_ARG, "a field array needs a .bb path");
    MappedFile f(mesh_path);
    const char* d = f.data();
    const size_t n = f.size();
    const int64_t NV = nv, NT = nt;
    std::memset(tags, 0, (size_t)nv);

    size_t p = find_key(d, n, 0, "Vertices");
    int64_t count = -1;
    TRN_REQUIRE(p < n && read_count(d, n, p, count) && count == NV, TRN_ERR_ARG,
                "Vertices count differs from nv");
    p = parse_section(d, n, records_start(d, n, p, "Vertices"), count, 3, 2, "Vertices", [&](int64_t k, const Fields& r) {
      xy[2 * k] = r.f[0];
      xy[2 * k + 1] = r.f[1];
      return true;
    });
    // Edges: both ends boundary (io.cpp:80-86)
    p = find_key(d, n, p, "Edges");
    count = 0;
    if (p < n && read_count(d, n, p, count))
      p = parse_section(d, n, records_start(d, n, p, "Edges"), count, 3, 0, "Edges", [&](int64_t, const Fields& r) {
        if (r.i[0] < 1 || r.i[0] > NV || r.i[1] < 1 || r.i[1] > NV) return false;
        tags[r.i[0] - 1] = kTagBound;
        tags[r.i[1] - 1] = kTagBound;
        return true;
      });
    p = find_key(d, n, p, "Triangles");
    count = -1;
    TRN_REQUIRE(p < n && read_count(d, n, p, count) && count == NT, TRN_ERR_ARG,
                "Triangles count differs from nt");
    p = parse_section(d, n, records_start(d, n, p, "Triangles"), count, 4, 0, "Triangles", [&](int64_t k, const Fields& r) {
      for (int j = 0; j < 3; ++j) {
        if (r.i[j] < 1 || r.i[j] > NV) return false;
        tri[3 * k + j] = (int32_t)(r.i[j] - 1);  // io.cpp:97-99
      }
      return true;
    });
    // Corners: boundary | corner, after the edges (io.cpp:101-104)
    p = find_key(d, n, p, "Corners");
    count = 0;
    if (p < n && read_count(d, n, p, count))
      parse_section(d, n, records_start(d, n, p, "Corners"), count, 1, 0, "Corners", [&](int64_t, const Fields& r) {
        if (r.i[0] < 1 || r.i[0] > NV) return false;
        tags[r.i[0] - 1] = kTagBound | kTagCorner;
        return true;
      });
    if (solut) {  // .bb: values from the second line on (io.cpp:108-112)
      MappedFile b(bb_path);
      const size_t q = next_line(b.data(), b.size(), 0);
      parse_section(b.data(), b.size(), q, NV, 1, 1, "solution", [&](int64_t k, const Fields& r) {
        solut[k] = r.f[0];
        return true;
      });
    }
  });
}

int trn_trnm_info(const char* path, int32_t* nv, int32_t* nt) {
  return guard_impl("trn_trnm_info", [&] {
    TRN_REQUIRE(path && nv && nt, TRN_ERR_ARG, "null argument");
    FILE* fp = std::fopen(path, "rb");
    TRN_REQUIRE(fp != nullptr, TRN_ERR_ARG, std::string("cannot open ") + path);
    char magic[4];
    int32_t h[3];
    const bool ok = std::fread(magic, 1, 4, fp) == 4 && std::fread(h, 4, 3, fp) == 3;
    std::fclose(fp);
    TRN_REQUIRE(ok && std::memcmp(magic, "TRNM", 4) == 0, TRN_ERR_ARG, "not a TRNM file");
    TRN_REQUIRE(h[0] == 1, TRN_ERR_ARG, "unsupported TRNM version");
    TRN_REQUIRE(h[1] > 0 && h[2] > 0, TRN_ERR_ARG, "empty TRNM mesh");
    *nv = h[1];
    *nt = h[2];
  });
}

int trn_trnm_read(trn_ctx* ctx, const char* path, int32_t nv, int32_t nt, double* xy,
                  int32_t* tri, uint8_t* tags, double* solut) {
  return guard_impl("trn_trnm_read", [&] {
    TRN_REQUIRE(xy && tri && tags, TRN_ERR_ARG, "null output array");
    int32_t fnv = 0, fnt = 0;
    const int rc = trn_trnm_info(path, &fnv, &fnt);
    TRN_REQUIRE(rc == TRN_OK, rc, trn_last_error());
    TRN_REQUIRE(fnv == nv && fnt == nt, TRN_ERR_ARG, "nv or nt differs from the file's");
    const int fd = ::open(path, O_RDONLY);
    TRN_REQUIRE(fd >= 0, TRN_ERR_ARG, std::string("cannot open ") + path);
    struct Part {
      void* dst;
      size_t bytes;
      off_t off;
    };
    const off_t o_xy = 16, o_tri = o_xy + (off_t)(16 * (size_t)nv),
                o_tags = o_tri + (off_t)(12 * (size_t)nt), o_sol = o_tags + (off_t)nv;
    const Part parts[4] = {{xy, 16 * (size_t)nv, o_xy},
                           {tri, 12 * (size_t)nt, o_tri},
                           {tags, (size_t)nv, o_tags},
                           {solut, 8 * (size_t)nv, o_sol}};
    const int T = io_threads();
    // pread of [o, o + len) into dst by T threads
    auto pread_all = [&](char* dst, off_t o, size_t len) -> bool {
      const size_t piece = std::max<size_t>(len / (size_t)T + 1, 1 << 20);
      std::atomic<bool> ok{true};
      std::vector<std::thread> th;
      for (size_t s = 0; s < len; s += piece) {
        th.emplace_back([&, s] {
          size_t done = 0, want = std::min(piece, len - s);
          while (done < want) {
            const ssize_t r = ::pread(fd, dst + s + done, want - done, o + (off_t)(s + done));
            if (r <= 0) {
              ok = false;
              return;
            }
            done += (size_t)r;
          }
        });
      }
      for (auto& t : th) t.join();
      return ok.load();
    };
    bool ok = true;
    if (ctx == nullptr) {  // host arrays
      for (const Part& pt : parts)
        if (pt.dst) ok = ok && pread_all((char*)pt.dst, pt.off, pt.bytes);
    } else {  // device arrays: two pinned buffers, copies overlap the next read
      DeviceScope ds(ctx);
      constexpr size_t kStage = size_t(64) << 20;
      char* stage[2] = {nullptr, nullptr};
      hipEvent_t done[2] = {nullptr, nullptr};
      try {
        for (int b = 0; b < 2; ++b) {
          TRN_HIP(hipHostMalloc((void**)&stage[b], kStage, hipHostMallocDefault));
          TRN_HIP(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
        }
        int b = 0;
        for (const Part& pt : parts) {
          for (size_t s = 0; ok && pt.dst && s < pt.bytes; s += kStage) {
            const size_t len = std::min(kStage, pt.bytes - s);
            TRN_HIP(hipEventSynchronize(done[b]));  // the buffer's last copy has finished
            ok = pread_all(stage[b], pt.off + (off_t)s, len);
            if (!ok) break;
            TRN_HIP(hipMemcpyAsync((char*)pt.dst + s, stage[b], len, hipMemcpyDefault,
                                   ctx->stream));
            TRN_HIP(hipEventRecord(done[b], ctx->stream));
            b ^= 1;
          }
        }
        TRN_HIP(hipStreamSynchronize(ctx->stream));
      } catch (...) {
        (void)hipStreamSynchronize(ctx->stream);
        for (int k = 0; k < 2; ++k) {
          if (stage[k]) (void)hipHostFree(stage[k]);
          if (done[k]) (void)hipEventDestroy(done[k]);
        }
        ::close(fd);
        throw;
      }
      for (int k = 0; k < 2; ++k) {
        (void)hipHostFree(stage[k]);
        (void)hipEventDestroy(done[k]);
      }
    }
    ::close(fd);
    TRN_REQUIRE(ok, TRN_ERR_ARG, "TRNM file shorter than its header says");
  });
}

int trn_trnm_write(const char* path, int32_t nv, int32_t nt, const double* xy,
                   const int32_t* tri, const uint8_t* tags, const double* solut) {
  return guard_impl("trn_trnm_write", [&] {
    TRN_REQUIRE(path && xy && tri && tags && solut, TRN_ERR_ARG, "null argument");
    TRN_REQUIRE(nv > 0 && nt > 0, TRN_ERR_ARG, "empty mesh");
    FILE* fp = std::fopen(path, "wb");
    TRN_REQUIRE(fp != nullptr, TRN_ERR_ARG, std::string("cannot create ") + path);
    const int32_t h[3] = {1, nv, nt};
    bool ok = std::fwrite("TRNM", 1, 4, fp) == 4 && std::fwrite(h, 4, 3, fp) == 3 &&
              std::fwrite(xy, 16, (size_t)nv, fp) == (size_t)nv &&
              std::fwrite(tri, 12, (size_t)nt, fp) == (size_t)nt &&
              std::fwrite(tags, 1, (size_t)nv, fp) == (size_t)nv &&
              std::fwrite(solut, 8, (size_t)nv, fp) == (size_t)nv;
    ok = (std::fclose(fp) == 0) && ok;
    TRN_REQUIRE(ok, TRN_ERR_ARG, std::string("cannot write ") + path);
  });
}

}  // extern "C"
