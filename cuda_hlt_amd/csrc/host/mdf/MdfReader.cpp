// MDF reading into the per-type raw-event layout (mdf/src/read_mdf.cpp:28-337,
// mdf/src/raw_helpers.cpp:8-21, mdf/src/compression.cpp:51-130).
//
// An MDF record is a header (three copies of the record size, checksum,
// compression byte, header-version/subheader-length byte, data type),
// a subheader of (hdr & 0xF) 32-bit words, one pad byte, then the raw banks,
// possibly compressed as a sequence of ROOT "ZL" blocks.  Each raw bank is
// u16 magic 0xCBCB, u16 length (8-byte bank header + data), u8 type,
// u8 version, u16 source id, data, padded to 4 bytes.
#include "velo_mdf.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr size_t kHeaderFixed = 19;  // sizes[3], checksum, compression, hdr, data type
constexpr uint16_t kBankMagic = 0xCBCB;

uint32_t rd32(const unsigned char* p)
{
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}
uint16_t rd16(const unsigned char* p)
{
  uint16_t v;
  std::memcpy(&v, p, 2);
  return v;
}

// ROOT zip blocks: 'Z' 'L' method, 3-byte compressed size, 3-byte
// uncompressed size, then a zlib stream of the compressed size.
std::vector<unsigned char> inflate_blocks(const unsigned char* src, size_t len)
{
  std::vector<unsigned char> out;
  size_t p = 0;
  while (p + 9 <= len) {
    if (src[p] != 'Z' || src[p + 1] != 'L' || src[p + 2] != Z_DEFLATED)
      throw std::runtime_error("unsupported compression block (not a zlib 'ZL' block)");
    const size_t csize = src[p + 3] | (size_t) src[p + 4] << 8 | (size_t) src[p + 5] << 16;
    const size_t usize = src[p + 6] | (size_t) src[p + 7] << 8 | (size_t) src[p + 8] << 16;
    if (p + 9 + csize > len) throw std::runtime_error("truncated compression block");
    const size_t at = out.size();
    out.resize(at + usize);
    z_stream zs {};
    zs.next_in = const_cast<unsigned char*>(src + p + 9);
    zs.avail_in = (uInt) csize;
    zs.next_out = out.data() + at;
    zs.avail_out = (uInt) usize;
    if (inflateInit(&zs) != Z_OK) throw std::runtime_error("inflateInit failed");
    const int rc = inflate(&zs, Z_FINISH);
    inflateEnd(&zs);
    if (rc != Z_STREAM_END || zs.total_out != usize) throw std::runtime_error("corrupt zlib block");
    p += 9 + csize;
  }
  if (p != len) throw std::runtime_error("trailing bytes after the compression blocks");
  return out;
}

// Builds, for each requested bank type, the event blob made of the banks
// of that type in [banks, banks + len): u32 n_banks, u32 byte offsets
// [n_banks + 1], then per bank the source id and the bank's data words
// (read_mdf.cpp:131-183).  A first walk checks the banks and sizes the
// blobs, a second copies each bank's data at once.
std::vector<std::vector<char>> bank_events(const unsigned char* banks, size_t len, const uint32_t* types,
                                           uint32_t n_types)
{
  auto slot = [&](uint8_t type) {
    for (uint32_t t = 0; t < n_types; ++t)
      if (types[t] == type) return (int) t;
    return -1;
  };
  std::vector<uint32_t> n_banks(n_types, 0);
  std::vector<size_t> n_words(n_types, 0);
  size_t q = 0;
  while (q + 8 <= len) {
    const uint16_t length = rd16(banks + q + 2);
    if (rd16(banks + q) != kBankMagic) throw std::runtime_error("raw bank with a bad magic pattern");
    if (length < 8 || q + length > len) throw std::runtime_error("raw bank overruns its record");
    const int t = slot(banks[q + 4]);
    if (t >= 0) {
      ++n_banks[t];
      n_words[t] += 1 + (length - 8) / 4;
    }
    q += (length + 3u) & ~3u;
  }
  std::vector<std::vector<char>> out(n_types);
  std::vector<uint32_t> at(n_types, 0), b(n_types, 0);
  for (uint32_t t = 0; t < n_types; ++t) {
    out[t].resize(4 * (2 + (size_t) n_banks[t] + n_words[t]));
    std::memcpy(out[t].data(), &n_banks[t], 4);
    std::memcpy(out[t].data() + 4, &at[t], 4);
  }
  for (q = 0; q + 8 <= len; q += (rd16(banks + q + 2) + 3u) & ~3u) {
    const int t = slot(banks[q + 4]);
    if (t < 0) continue;
    char* offsets = out[t].data() + 4;
    char* data = out[t].data() + 4 * (2 + (size_t) n_banks[t]);
    const uint32_t source = rd16(banks + q + 6);
    const size_t bytes = 4 * ((rd16(banks + q + 2) - 8) / 4);
    std::memcpy(data + at[t], &source, 4);
    std::memcpy(data + at[t] + 4, banks + q + 8, bytes);
    at[t] += (uint32_t) (4 + bytes);
    std::memcpy(offsets + 4 * (++b[t]), &at[t], 4);
  }
  return out;
}

struct Record {
  const unsigned char* bytes; // into the file's mapping
  size_t size;
  uint32_t index;             // position among all records read
};

// Checksum, decompression and bank walk of one whole record.
std::vector<std::vector<char>> decode_record(const Record& r, bool verify_checksum, const uint32_t* types,
                                             uint32_t n_types)
{
  const unsigned char* rec = r.bytes;
  const size_t size = r.size;
  const uint32_t checksum = rd32(rec + 12);
  if (verify_checksum && checksum && velo_mdf_checksum(rec + 16, size - 16) != checksum)
    throw std::runtime_error("MDF record " + std::to_string(r.index) + ": checksum mismatch");
  const size_t start = kHeaderFixed + (size_t) (rec[17] & 0x0F) * 4 + 1; // pad byte after the subheader
  if (rec[16] & 0x0F) {
    const std::vector<unsigned char> banks = inflate_blocks(rec + start, size - start);
    return bank_events(banks.data(), banks.size(), types, n_types);
  }
  return bank_events(rec + start, size - start, types, n_types);
}

// A read-only mapping of one file.
struct Mapping {
  const unsigned char* data = nullptr;
  size_t size = 0;
  explicit Mapping(const char* path)
  {
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) throw std::runtime_error(std::string("cannot open ") + path);
    struct stat st {};
    if (::fstat(fd, &st) != 0) {
      ::close(fd);
      throw std::runtime_error(std::string("cannot stat ") + path);
    }
    size = (size_t) st.st_size;
    if (size) {
      void* p = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (p == MAP_FAILED) {
        ::close(fd);
        throw std::runtime_error(std::string("cannot map ") + path);
      }
      ::madvise(p, size, MADV_SEQUENTIAL);
      data = static_cast<const unsigned char*>(p);
    }
    ::close(fd);
  }
  ~Mapping()
  {
    if (data) ::munmap(const_cast<unsigned char*>(data), size);
  }
  Mapping(const Mapping&) = delete;
  Mapping& operator=(const Mapping&) = delete;
};

// Finds the records [start_event, start_event + max_events) of the files.
std::vector<Record> find_records(const std::vector<std::unique_ptr<Mapping>>& files, uint32_t max_events,
                                 uint32_t start_event, uint32_t& n_records)
{
  std::vector<Record> out;
  n_records = 0;
  for (const auto& f : files) {
    size_t at = 0;
    while ((max_events == 0 || out.size() < max_events) && at < f->size) {
      const unsigned char* head = f->data + at;
      if (f->size - at < kHeaderFixed) throw std::runtime_error("truncated MDF header");
      const uint32_t size = rd32(head);
      if ((head[17] >> 4) != 3) throw std::runtime_error("unsupported MDF header version (only 3 is defined)");
      if (size < kHeaderFixed + (size_t) (head[17] & 0x0F) * 4 + 1)
        throw std::runtime_error("MDF record shorter than its header");
      if (size > f->size - at) throw std::runtime_error("truncated MDF record");
      if (n_records >= start_event) out.push_back(Record {head, size, n_records});
      ++n_records;
      at += size;
    }
  }
  return out;
}

} // namespace

extern "C" {

uint32_t velo_mdf_checksum(const void* data, size_t len)
{
  // one-at-a-time hash; bytes are signed as in the reference
  const signed char* k = static_cast<const signed char*>(data);
  uint32_t hash = 0;
  for (size_t i = 0; i < len; ++i) {
    hash += (uint32_t) (int32_t) k[i];
    hash += hash << 10;
    hash ^= hash >> 6;
  }
  hash += hash << 3;
  hash ^= hash >> 11;
  hash += hash << 15;
  return hash;
}

int velo_mdf_read_banks(const char* const* paths, uint32_t n_paths, uint32_t max_events, uint32_t start_event,
                        int verify_checksum, uint32_t n_threads, const uint32_t* bank_types, uint32_t n_types,
                        velo_mdf_events* outs, char* err, size_t err_len)
{
  auto fail = [&](const std::string& msg) {
    if (err && err_len) std::snprintf(err, err_len, "%s", msg.c_str());
    return -1;
  };
  if (!outs || !bank_types || n_types == 0 || (n_paths && !paths)) return fail("velo_mdf_read_banks: null argument");
  for (uint32_t t = 0; t < n_types; ++t) std::memset(&outs[t], 0, sizeof(outs[t]));
  for (uint32_t t = 0; t < n_types; ++t) {
    if (bank_types[t] > 255) return fail("bank type " + std::to_string(bank_types[t]) + " is not a raw bank type");
    for (uint32_t u = 0; u < t; ++u)
      if (bank_types[u] == bank_types[t]) return fail("bank type " + std::to_string(bank_types[t]) + " requested twice");
  }
  std::vector<std::unique_ptr<Mapping>> files;
  std::vector<Record> records;
  uint32_t n_records = 0;
  try {
    for (uint32_t f = 0; f < n_paths; ++f) files.push_back(std::make_unique<Mapping>(paths[f]));
    records = find_records(files, max_events, start_event, n_records);
  } catch (const std::exception& e) {
    return fail(e.what());
  }

  // Records decode independently: thread t takes records t, t + T, ...
  const size_t n = records.size();
  std::vector<std::vector<std::vector<char>>> events(n);
  const size_t T = std::max<size_t>(1, std::min<size_t>(n_threads ? n_threads : 1, n));
  auto parallel = [T](auto&& body) {
    if (T == 1) return body(0);
    std::vector<std::thread> pool;
    for (size_t t = 0; t < T; ++t) pool.emplace_back(body, t);
    for (auto& th : pool) th.join();
  };
  std::vector<std::string> errors(T);
  parallel([&](size_t t) {
    try {
      for (size_t i = t; i < n; i += T)
        events[i] = decode_record(records[i], verify_checksum != 0, bank_types, n_types);
    } catch (const std::exception& e) {
      errors[t] = e.what();
    }
  });
  // the error of the earliest failing record is reported
  size_t first_bad = n;
  for (size_t t = 0; t < T; ++t)
    if (!errors[t].empty())
      for (size_t i = t; i < n; i += T)
        if (events[i].empty()) {
          first_bad = std::min(first_bad, i);
          break;
        }
  if (first_bad < n) return fail(errors[first_bad % T]);

  for (uint32_t k = 0; k < n_types; ++k) {
    velo_mdf_events* out = &outs[k];
    std::vector<size_t> offsets(n + 1, 0);
    for (size_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + events[i][k].size();
    const size_t total = offsets[n];
    if (total > 0xFFFFFFFFull) {
      for (uint32_t u = 0; u < n_types; ++u) velo_mdf_free(&outs[u]);
      return fail("more than 4 GiB of events of bank type " + std::to_string(bank_types[k]) +
                  " (offsets are 32-bit)");
    }
    out->events = static_cast<char*>(std::malloc(total ? total : 1));
    out->offsets = static_cast<uint32_t*>(std::malloc(4 * (n + 1)));
    if (!out->events || !out->offsets) {
      for (uint32_t u = 0; u < n_types; ++u) velo_mdf_free(&outs[u]);
      return fail("out of host memory");
    }
    for (size_t i = 0; i <= n; ++i) out->offsets[i] = (uint32_t) offsets[i];
    parallel([&](size_t t) {
      for (size_t i = t; i < n; i += T) {
        std::memcpy(out->events + offsets[i], events[i][k].data(), events[i][k].size());
        std::vector<char>().swap(events[i][k]);
      }
    });
    out->n_events = (uint32_t) n;
    out->n_records = n_records;
    out->events_size = total;
  }
  return 0;
}

int velo_mdf_read_vp(const char* const* paths, uint32_t n_paths, uint32_t max_events, uint32_t start_event,
                     int verify_checksum, uint32_t n_threads, velo_mdf_events* out, char* err, size_t err_len)
{
  const uint32_t vp = VELO_MDF_BANK_VP;
  return velo_mdf_read_banks(paths, n_paths, max_events, start_event, verify_checksum, n_threads, &vp, 1, out, err,
                             err_len);
}

void velo_mdf_free(velo_mdf_events* ev)
{
  if (!ev) return;
  std::free(ev->events);
  std::free(ev->offsets);
  ev->events = nullptr;
  ev->offsets = nullptr;
  ev->events_size = 0;
  ev->n_events = 0;
}

} // extern "C"
