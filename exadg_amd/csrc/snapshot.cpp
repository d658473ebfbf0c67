// Snapshot I/O of solution vectors: the counterpart of ExaDG's restart files
// (time_integration/restart.h:48-83 file naming, rotation of old snapshots and writing;
// :194-300 the parameters a discretization is serialized with and their check on reading).
// A snapshot holds the discretization parameters, the time, the step and caller data, and a
// list of equally long vectors (device or host memory), streamed through a bounded pinned
// staging buffer, followed by a checksum over every byte before it.
//
// File layout (little endian): "EXADGSN1", then the header fields in the order of
// exadg_snapshot_info (each field at its natural width, no padding), then the vectors, then the
// FNV-1a 64-bit hash of everything before it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/exadg_amd.h"
#include "common.h"
#include "snapshot.h"

namespace ExaDG
{
namespace
{
constexpr char     kMagic[8]    = {'E', 'X', 'A', 'D', 'G', 'S', 'N', '1'};
constexpr int32_t  kVersion     = 1;
constexpr size_t   kStageBytes  = size_t(64) << 20;  // pinned staging buffer
constexpr uint64_t kFnvOffset   = 0xcbf29ce484222325ull;
constexpr uint64_t kFnvPrime    = 0x100000001b3ull;

struct Fnv
{
  uint64_t h = kFnvOffset;
  void     add(void const * p, size_t n)
  {
    auto const * b = static_cast<unsigned char const *>(p);
    for(size_t i = 0; i < n; ++i)
      h = (h ^ b[i]) * kFnvPrime;
  }
};

struct File
{
  FILE * f = nullptr;
  ~File()
  {
    if(f)
      std::fclose(f);
  }
};

// header fields in file order
template<typename F>
void for_each_field(exadg_snapshot_info & s, F && f)
{
  f(&s.version, sizeof(s.version));
  f(&s.precision, sizeof(s.precision));
  f(&s.n_vectors, sizeof(s.n_vectors));
  f(&s.n, sizeof(s.n));
  f(&s.degree, sizeof(s.degree));
  f(s.mesh.n_cells, sizeof(s.mesh.n_cells));
  f(s.mesh.lower, sizeof(s.mesh.lower));
  f(s.mesh.upper, sizeof(s.mesh.upper));
  f(&s.mesh.deformation, sizeof(s.mesh.deformation));
  f(&s.mesh.frequency, sizeof(s.mesh.frequency));
  f(&s.mesh.mapping_degree, sizeof(s.mesh.mapping_degree));
  f(s.mesh.periodic, sizeof(s.mesh.periodic));
  f(&s.time, sizeof(s.time));
  f(&s.step, sizeof(s.step));
  f(s.user, sizeof(s.user));
}

size_t scalar_size(int precision)
{
  if(precision == EXADG_FP64)
    return 8;
  if(precision == EXADG_FP32)
    return 4;
  EXADG_THROW(EXADG_ERR_ARGUMENT, "snapshot precision must be EXADG_FP64 or EXADG_FP32");
}

// pinned staging for device vectors (freed on scope exit)
struct Stage
{
  void * p = nullptr;
  explicit Stage(bool on_device)
  {
    if(on_device)
      HIP_CHECK(hipHostMalloc(&p, kStageBytes, hipHostMallocDefault));
  }
  ~Stage()
  {
    if(p)
      (void)hipHostFree(p);
  }
};

bool file_exists(std::string const & name)
{
  FILE * f = std::fopen(name.c_str(), "rb");
  if(f)
    std::fclose(f);
  return f != nullptr;
}

void read_exact(FILE * f, void * p, size_t n, Fnv & h, std::string const & path)
{
  if(n > 0 && std::fread(p, 1, n, f) != n)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "snapshot " + path + " is truncated");
  h.add(p, n);
}

exadg_snapshot_info read_header(FILE * f, Fnv & h, std::string const & path)
{
  char magic[8];
  read_exact(f, magic, 8, h, path);
  if(std::memcmp(magic, kMagic, 8) != 0)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "not a snapshot file: " + path);
  exadg_snapshot_info s;
  std::memset(&s, 0, sizeof(s));
  for_each_field(s, [&](void * p, size_t n) { read_exact(f, p, n, h, path); });
  if(s.version != kVersion)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "snapshot " + path + " has an unsupported version");
  if(s.n_vectors < 0 || s.n < 0 || (s.precision != EXADG_FP64 && s.precision != EXADG_FP32))
    EXADG_THROW(EXADG_ERR_ARGUMENT, "snapshot " + path + " has an invalid header");
  return s;
}
}  // namespace

void rotate_snapshots(std::string const & path, int n_keep)
{
  // keep n_keep snapshots: path -> path.old.1 -> ... -> path.old.{n_keep-1} (restart.h:57-72)
  for(int i = n_keep - 2; i >= 0; --i)
  {
    std::string const from = i == 0 ? path : path + ".old." + std::to_string(i);
    std::string const to   = path + ".old." + std::to_string(i + 1);
    if(file_exists(from) && std::rename(from.c_str(), to.c_str()) != 0)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "can not rename " + from + " -> " + to);
  }
}

void write_snapshot(std::string const & path, exadg_snapshot_info const & info_in,
                    void const * const * vectors, bool on_device, int n_keep, hipStream_t s)
{
  exadg_snapshot_info info = info_in;
  info.version             = kVersion;
  size_t const sz          = scalar_size(info.precision);
  if(info.n_vectors < 0 || info.n < 0 || (info.n_vectors > 0 && vectors == nullptr))
    EXADG_THROW(EXADG_ERR_ARGUMENT, "invalid snapshot vectors");
  for(int v = 0; v < info.n_vectors; ++v)
    if(vectors[v] == nullptr && info.n > 0)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "null snapshot vector");
  if(n_keep < 1)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "n_snapshots_keep must be at least 1");
  // write to a temporary name first, so that a failed write leaves the old snapshots intact
  std::string const tmp = path + ".tmp";
  Fnv               h;
  {
    File out;
    out.f = std::fopen(tmp.c_str(), "wb");
    if(!out.f)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "can not open " + tmp + " for writing");
    auto put = [&](void const * p, size_t n) {
      if(n > 0 && std::fwrite(p, 1, n, out.f) != n)
        EXADG_THROW(EXADG_ERR_ARGUMENT, "write to " + tmp + " failed");
      h.add(p, n);
    };
    put(kMagic, 8);
    for_each_field(info, [&](void * p, size_t n) { put(p, n); });
    Stage        stage(on_device);
    size_t const bytes = (size_t)info.n * sz;
    for(int v = 0; v < info.n_vectors; ++v)
    {
      auto const * src = static_cast<char const *>(vectors[v]);
      for(size_t off = 0; off < bytes; off += kStageBytes)
      {
        size_t const len = std::min(kStageBytes, bytes - off);
        if(on_device)
        {
          HIP_CHECK(hipMemcpyAsync(stage.p, src + off, len, hipMemcpyDeviceToHost, s));
          HIP_CHECK(hipStreamSynchronize(s));
          put(stage.p, len);
        }
        else
          put(src + off, len);
      }
    }
    uint64_t const sum = h.h;
    if(std::fwrite(&sum, 1, 8, out.f) != 8 || std::fflush(out.f) != 0)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "write to " + tmp + " failed");
  }
  rotate_snapshots(path, n_keep);
  if(std::rename(tmp.c_str(), path.c_str()) != 0)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "can not rename " + tmp + " -> " + path);
}

exadg_snapshot_info read_snapshot_info(std::string const & path)
{
  File in;
  in.f = std::fopen(path.c_str(), "rb");
  if(!in.f)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "can not open snapshot " + path);
  Fnv h;
  return read_header(in.f, h, path);
}

exadg_snapshot_info read_snapshot(std::string const & path, void * const * vectors, int n_vectors,
                                  int64_t n, int precision, bool on_device, hipStream_t s)
{
  File in;
  in.f = std::fopen(path.c_str(), "rb");
  if(!in.f)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "can not open snapshot " + path);
  Fnv                       h;
  exadg_snapshot_info const info = read_header(in.f, h, path);
  // the caller states what it expects, as read_deserialization_parameters is checked against
  // the running discretization
  if(info.n_vectors != n_vectors || info.n != n || info.precision != precision)
    EXADG_THROW(EXADG_ERR_ARGUMENT,
                "snapshot " + path + " holds " + std::to_string(info.n_vectors) + " vectors of " +
                  std::to_string(info.n) + " entries (precision " + std::to_string(info.precision) +
                  "), not " + std::to_string(n_vectors) + " of " + std::to_string(n) +
                  " (precision " + std::to_string(precision) + ")");
  if(n_vectors > 0 && vectors == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector list");
  for(int v = 0; v < n_vectors; ++v)
    if(vectors[v] == nullptr && n > 0)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "null snapshot vector");
  size_t const sz    = scalar_size(precision);
  size_t const bytes = (size_t)n * sz;
  // two passes through a bounded staging buffer: verify the checksum over the whole file first,
  // then copy, so that no destination is written from a corrupted snapshot
  long const data_begin = std::ftell(in.f);
  Stage      stage(on_device);
  std::unique_ptr<char[]> plain;
  char * buf = static_cast<char *>(stage.p);
  if(!buf)
  {
    plain.reset(new char[kStageBytes]);
    buf = plain.get();
  }
  size_t const total = bytes * (size_t)n_vectors;
  for(size_t off = 0; off < total; off += kStageBytes)
    read_exact(in.f, buf, std::min(kStageBytes, total - off), h, path);
  uint64_t sum = 0;
  if(std::fread(&sum, 1, 8, in.f) != 8)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "snapshot " + path + " is truncated");
  if(sum != h.h)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "snapshot " + path + " fails its checksum (corrupted)");
  if(std::fseek(in.f, data_begin, SEEK_SET) != 0)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "can not seek in snapshot " + path);
  Fnv unused;
  for(int v = 0; v < n_vectors; ++v)
    for(size_t off = 0; off < bytes; off += kStageBytes)
    {
      size_t const len = std::min(kStageBytes, bytes - off);
      char *       dst = static_cast<char *>(vectors[v]) + off;
      if(on_device)
      {
        read_exact(in.f, buf, len, unused, path);
        HIP_CHECK(hipMemcpyAsync(dst, buf, len, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
      }
      else
        read_exact(in.f, dst, len, unused, path);
    }
  return info;
}
}  // namespace ExaDG
