// Per-event input files in a folder (main/src/InputTools.cpp:118-211:
// list_folder, natural_order, read_folder; main/src/CheckVP.cpp:5-52:
// check_velopix_events) and the MC truth of the checker
// (checker/tracking/src/MCEvent.cpp:16-103,
// checker/tracking/src/CheckerInvoker.cpp:3-44).
//
// Folder contents are file names "N.bin" or "N_M.bin" (N, M decimal),
// ordered as the pairs (0, N) and (N, M) compare as numbers.  Event files
// are read whole and concatenated; the event layout is whatever the files
// hold (a VP event: u32 n_banks, u32 offsets[n_banks + 1], banks).
#include <dirent.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "velo_checker.h"
#include "velo_mdf.h"

namespace {

// (0, N) for "N.bin", (N, M) for "N_M.bin"; false for any other name.
bool bin_number(const std::string& name, std::pair<long long, long long>* key)
{
  if (name.size() < 5 || name.compare(name.size() - 4, 4, ".bin") != 0) return false;
  const std::string stem = name.substr(0, name.size() - 4);
  const size_t us = stem.find('_');
  auto digits = [](const std::string& s) {
    return !s.empty() && s.size() <= 18 && std::all_of(s.begin(), s.end(), [](char c) { return c >= '0' && c <= '9'; });
  };
  if (us == std::string::npos) {
    if (!digits(stem)) return false;
    *key = {0, std::stoll(stem)};
    return true;
  }
  const std::string a = stem.substr(0, us), b = stem.substr(us + 1);
  if (!digits(a) || !digits(b)) return false;
  *key = {std::stoll(a), std::stoll(b)};
  return true;
}

// Sorted event file names of a folder; throws on a missing or empty folder,
// or on a name that is not an event file.
std::vector<std::string> list_event_files(const std::string& folder)
{
  DIR* dir = opendir(folder.c_str());
  if (!dir) throw std::runtime_error("folder " + folder + " could not be opened");
  std::vector<std::pair<std::pair<long long, long long>, std::string>> files;
  std::string bad;
  while (dirent* ent = readdir(dir)) {
    const std::string name = ent->d_name;
    if (name == "." || name == "..") continue;
    std::pair<long long, long long> key;
    if (!bin_number(name, &key)) {
      if (bad.empty()) bad = name;
      continue;
    }
    files.emplace_back(key, name);
  }
  closedir(dir);
  if (!bad.empty()) throw std::runtime_error("folder " + folder + " holds " + bad + ", not an N.bin / N_M.bin file");
  if (files.empty()) throw std::runtime_error("no .bin files found in folder " + folder);
  std::stable_sort(files.begin(), files.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<std::string> out;
  out.reserve(files.size());
  for (auto& f : files) out.push_back(std::move(f.second));
  return out;
}

std::vector<char> read_file(const std::string& path)
{
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error("file " + path + " could not be read");
  const std::streamsize size = in.tellg();
  in.seekg(0);
  std::vector<char> data((size_t) size);
  if (size > 0 && !in.read(data.data(), size)) throw std::runtime_error("file " + path + " could not be read");
  return data;
}

void set_error(char* err, size_t len, const std::string& msg)
{
  if (err && len) std::snprintf(err, len, "%s", msg.c_str());
}

template<typename T>
T take(const char*& p, const char* end)
{
  if ((size_t)(end - p) < sizeof(T)) throw std::runtime_error("event ends inside a record");
  T v;
  std::memcpy(&v, p, sizeof(T));
  p += sizeof(T);
  return v;
}

} // namespace

extern "C" {

int velo_read_folder(const char* folder, uint32_t n_events, uint32_t start_event, velo_mdf_events* out, char* err,
                     size_t err_len)
{
  if (!folder || !out) {
    set_error(err, err_len, "velo_read_folder: null argument");
    return -1;
  }
  std::memset(out, 0, sizeof(*out));
  try {
    const std::vector<std::string> files = list_event_files(folder);
    const size_t n = n_events ? n_events : files.size();
    std::vector<std::vector<char>> data(n);
    size_t total = 0;
    for (size_t i = 0; i < n; ++i) {
      // events past the last file start again from the first (round-robin)
      data[i] = read_file(std::string(folder) + "/" + files[(start_event + i) % files.size()]);
      total += data[i].size();
    }
    if (total > 0xFFFFFFFFull) throw std::runtime_error("more than 4 GiB of events (offsets are 32-bit)");
    out->events = static_cast<char*>(std::malloc(total ? total : 1));
    out->offsets = static_cast<uint32_t*>(std::malloc(4 * (n + 1)));
    if (!out->events || !out->offsets) {
      velo_mdf_free(out);
      throw std::runtime_error("out of host memory");
    }
    size_t pos = 0;
    for (size_t i = 0; i < n; ++i) {
      out->offsets[i] = (uint32_t) pos;
      if (!data[i].empty()) std::memcpy(out->events + pos, data[i].data(), data[i].size());
      pos += data[i].size();
    }
    out->offsets[n] = (uint32_t) pos;
    out->events_size = total;
    out->n_events = (uint32_t) n;
    out->n_records = (uint32_t) files.size();
  }
  catch (const std::exception& e) {
    set_error(err, err_len, std::string("velo_read_folder: ") + e.what());
    return -1;
  }
  return 0;
}

int64_t velo_check_vp_events(const char* events, const uint32_t* offsets, uint32_t n_events)
{
  if (!events || !offsets) return -1;
  int64_t mismatches = 0;
  for (uint32_t e = 0; e < n_events; ++e) {
    if (offsets[e + 1] < offsets[e]) return -1;
    const size_t words = (offsets[e + 1] - offsets[e]) / 4;
    const char* base = events + offsets[e];
    auto word = [&](size_t i) {
      uint32_t v;
      std::memcpy(&v, base + 4 * i, 4);
      return v;
    };
    if (words < 1) return -1;
    const uint32_t n_banks = word(0);
    if ((size_t) n_banks + 2 > words) return -1;
    const size_t payload = 2 + (size_t) n_banks;
    for (uint32_t b = 0; b < n_banks; ++b) {
      const uint32_t off = word(1 + b);
      if (off % 4 || payload + off / 4 >= words) return -1;
      if (word(payload + off / 4) != b) ++mismatches;
    }
  }
  return mismatches;
}

int velo_mc_read_folder(const char* folder, uint32_t n_events, uint32_t start_event, velo_mc_truth* out, char* err,
                        size_t err_len)
{
  if (!folder || !out) {
    set_error(err, err_len, "velo_mc_read_folder: null argument");
    return -1;
  }
  std::memset(out, 0, sizeof(*out));
  try {
    const std::vector<std::string> files = list_event_files(folder);
    const size_t n = n_events ? n_events : files.size();
    if (start_event + n > files.size())
      throw std::runtime_error("requested " + std::to_string(n) + " events from " + std::to_string(start_event) +
                               ", but only " + std::to_string(files.size()) + " MC files are present");
    std::vector<velo_gen_mcp> mcps;
    std::vector<uint32_t> hits;
    for (size_t i = 0; i < n; ++i) {
      const std::string path = std::string(folder) + "/" + files[start_event + i];
      const std::vector<char> data = read_file(path);
      const char* p = data.data();
      const char* end = p + data.size();
      try {
        const uint32_t n_mcp = take<uint32_t>(p, end);
        for (uint32_t k = 0; k < n_mcp; ++k) {
          velo_gen_mcp m {};
          m.event = (uint32_t) i;
          m.key = take<uint32_t>(p, end);
          m.pid = take<int32_t>(p, end);
          m.p = take<float>(p, end);
          m.pt = take<float>(p, end);
          m.eta = take<float>(p, end);
          m.phi = take<float>(p, end);
          const int8_t is_long = take<int8_t>(p, end);
          (void) take<int8_t>(p, end); // isDown: no VELO category reads it
          const int8_t has_velo = take<int8_t>(p, end);
          (void) take<int8_t>(p, end); // hasUT
          (void) take<int8_t>(p, end); // hasSciFi
          const int8_t from_b = take<int8_t>(p, end);
          const int8_t from_d = take<int8_t>(p, end);
          const int8_t from_s = take<int8_t>(p, end);
          m.n_pv = take<uint32_t>(p, end);
          m.flags = (has_velo ? VELO_MCP_HAS_VELO : 0u) | (is_long ? VELO_MCP_IS_LONG : 0u) |
                    (m.pid == 11 || m.pid == -11 ? VELO_MCP_IS_ELECTRON : 0u) | (m.eta < 0.f ? VELO_MCP_BACKWARD : 0u) |
                    (from_b ? VELO_MCP_FROM_B : 0u) | (from_d ? VELO_MCP_FROM_D : 0u) |
                    (from_s ? VELO_MCP_FROM_S : 0u);
          // VELO hits, then the UT and SciFi hits (skipped); a particle is
          // listed for the VELO check only with VELO hits (MCEvent.cpp:57-64)
          const uint32_t n_velo = take<uint32_t>(p, end);
          m.hit_offset = (uint32_t) hits.size();
          m.n_hits = n_velo;
          for (uint32_t h = 0; h < n_velo; ++h) hits.push_back(take<uint32_t>(p, end));
          for (int det = 0; det < 2; ++det) {
            const uint32_t n_other = take<uint32_t>(p, end);
            if ((size_t)(end - p) / 4 < n_other) throw std::runtime_error("event ends inside a hit list");
            p += 4 * (size_t) n_other;
          }
          if (n_velo > 0) mcps.push_back(m);
          else hits.resize(m.hit_offset);
        }
        if (p != end)
          throw std::runtime_error("size mismatch: " + std::to_string(p - data.data()) + " bytes read of " +
                                   std::to_string(data.size()));
      }
      catch (const std::exception& e) {
        throw std::runtime_error(path + ": " + e.what());
      }
    }
    out->mcps = static_cast<velo_gen_mcp*>(std::malloc(sizeof(velo_gen_mcp) * (mcps.size() ? mcps.size() : 1)));
    out->mc_hits = static_cast<uint32_t*>(std::malloc(4 * (hits.size() ? hits.size() : 1)));
    if (!out->mcps || !out->mc_hits) {
      velo_mc_free(out);
      throw std::runtime_error("out of host memory");
    }
    if (!mcps.empty()) std::memcpy(out->mcps, mcps.data(), sizeof(velo_gen_mcp) * mcps.size());
    if (!hits.empty()) std::memcpy(out->mc_hits, hits.data(), 4 * hits.size());
    out->n_mcps = mcps.size();
    out->n_mc_hits = hits.size();
    out->n_events = (uint32_t) n;
  }
  catch (const std::exception& e) {
    set_error(err, err_len, std::string("velo_mc_read_folder: ") + e.what());
    return -1;
  }
  return 0;
}

void velo_mc_free(velo_mc_truth* t)
{
  if (!t) return;
  std::free(t->mcps);
  std::free(t->mc_hits);
  std::memset(t, 0, sizeof(*t));
}

} // extern "C"
