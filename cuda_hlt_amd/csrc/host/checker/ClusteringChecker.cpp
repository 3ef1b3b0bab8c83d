// Clustering checker (host): the reference's x86 clustering of raw VP banks
// (x86/velo/clustering/src/Clustering.cpp:3-270) as the expected clusters,
// and its efficiency / clone / ghost report
// (checker/clustering/src/ClusteringChecker.cpp:32-135).
#include "velo_checker.h"
#include "velo_hlt.h"

#include <cmath>
#include <cstring>
#include <set>
#include <vector>

namespace {

constexpr uint32_t kRows = 256;          // VP::NRows
constexpr uint32_t kColumns = 768;       // VP::NSensorColumns
constexpr uint32_t kChipColumns = 256;   // VP::ChipColumns

uint32_t rd32(const char* p)
{
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}

uint32_t cluster_id(uint32_t sensor, uint32_t cx, uint32_t cy)
{
  return 0x80000000u + ((sensor << 18) | ((cx / kChipColumns) << 16) | ((cx % kChipColumns) << 8) | cy);
}

// The x86 clustering of one sensor bank: isolated SPs through the pattern
// table (one or two clusters), the pixels of the other SPs through a
// row-scan flood fill seeded in the order the pixels were recorded
// (Clustering.cpp:54-262).  Appends the clusters' LHCb IDs.
class SensorClustering {
public:
  SensorClustering() : buffer_(kRows * kColumns, 0)
  {
    velo_sp_tables(patterns_, sizes_, fx_, fy_);
  }

  void bank(uint32_t sensor, const char* words, uint32_t sp_count, std::vector<uint32_t>& ids)
  {
    pixels_.clear();
    for (uint32_t j = 0; j < sp_count; ++j) {
      const uint32_t w = rd32(words + 4 * (size_t) j);
      const uint32_t sp = w & 0xFFu;
      if (sp == 0) continue; // Clustering.cpp:57-59
      const uint32_t addr = (w & 0x007FFF00u) >> 8;
      const uint32_t sp_row = addr & 0x3Fu, sp_col = addr >> 6;
      if (w & 0x80000000u) {
        const uint32_t idx = patterns_[sp];
        ids.push_back(cluster_id(sensor, sp_col * 2 + ((idx >> 2) & 1u), sp_row * 4 + (idx & 3u)));
        if (idx & 8u) ids.push_back(cluster_id(sensor, sp_col * 2 + ((idx >> 6) & 1u), sp_row * 4 + ((idx >> 4) & 3u)));
        continue;
      }
      for (uint32_t shift = 0; shift < 8; ++shift) {
        if (!((sp >> shift) & 1u)) continue;
        const uint32_t row = sp_row * 4 + shift % 4, col = sp_col * 2 + shift / 4;
        if (row >= kRows || col >= kColumns) continue; // outside the sensor: not recordable
        const uint32_t idx = (col << 8) | row;
        buffer_[idx] = 1;
        pixels_.push_back(idx);
      }
    }
    for (const uint32_t seed : pixels_) {
      if (!buffer_[seed]) continue; // taken by an earlier cluster
      buffer_[seed] = 0;
      stack_.assign(1, seed);
      uint32_t x = 0, y = 0, n = 0;
      while (!stack_.empty()) {
        const uint32_t idx = stack_.back();
        stack_.pop_back();
        const uint32_t row = idx & 0xFFu, col = idx >> 8;
        x += col;
        y += row;
        ++n;
        up_down(idx, row);
        for (uint32_t c = col + 1; c < kColumns; ++c) { // the row to the right
          const uint32_t nidx = (c << 8) | row;
          up_down(nidx, row);
          if (!buffer_[nidx]) break;
          buffer_[nidx] = 0;
          x += c;
          y += row;
          ++n;
        }
        for (int c = (int) col - 1; c >= 0; --c) { // the row to the left
          const uint32_t nidx = ((uint32_t) c << 8) | row;
          up_down(nidx, row);
          if (!buffer_[nidx]) break;
          buffer_[nidx] = 0;
          x += (uint32_t) c;
          y += row;
          ++n;
        }
      }
      ids.push_back(cluster_id(sensor, x / n, y / n));
    }
  }

private:
  // pixels above and below idx join the stack
  void up_down(uint32_t idx, uint32_t row)
  {
    if (row < kRows - 1 && buffer_[idx + 1]) {
      buffer_[idx + 1] = 0;
      stack_.push_back(idx + 1);
    }
    if (row > 0 && buffer_[idx - 1]) {
      buffer_[idx - 1] = 0;
      stack_.push_back(idx - 1);
    }
  }

  std::vector<uint8_t> buffer_;
  std::vector<uint32_t> pixels_, stack_;
  uint8_t patterns_[256], sizes_[256];
  float fx_[512], fy_[512];
};

// ClusteringChecker.cpp:73-88 / :98-113: the same ID, or (unless only IDs
// are compared) the same sensor and chip with a pixel distance below the
// allowed error.
bool matches(uint32_t a, uint32_t b, bool just_ids, float allowed)
{
  if (a == b) return true;
  if (just_ids || (a & 0xFFFF0000u) != (b & 0xFFFF0000u)) return false;
  const uint32_t dr = (a & 0xFFu) - (b & 0xFFu);
  const uint32_t dc = ((a >> 8) & 0xFFu) - ((b >> 8) & 0xFFu);
  const float d = (float) (dr * dr + dc * dc);
  return std::sqrt(d) < allowed;
}

} // namespace

extern "C" {

int velo_x86_clustering(const char* events, const uint32_t* offsets, uint32_t n_events, uint32_t* ids,
                        uint32_t* id_offsets, size_t capacity, size_t* n_ids)
{
  if (!events || !offsets || !id_offsets || !n_ids || (capacity && !ids)) return -1;
  SensorClustering sc;
  std::vector<uint32_t> ev_ids;
  size_t total = 0;
  id_offsets[0] = 0;
  for (uint32_t e = 0; e < n_events; ++e) {
    ev_ids.clear();
    const char* ev = events + offsets[e];
    const uint32_t n_banks = rd32(ev);
    const char* payload = ev + 4 * ((size_t) n_banks + 2);
    for (uint32_t b = 0; b < n_banks; ++b) {
      const char* bank = payload + rd32(ev + 4 * (1 + (size_t) b));
      const uint32_t sensor = rd32(bank);
      if (sensor >= VELO_N_SENSORS) return -1;
      sc.bank(sensor, bank + 8, rd32(bank + 4), ev_ids);
    }
    for (size_t i = 0; i < ev_ids.size(); ++i)
      if (total + i < capacity) ids[total + i] = ev_ids[i];
    total += ev_ids.size();
    id_offsets[e + 1] = (uint32_t) total;
  }
  *n_ids = total;
  return total > capacity ? 1 : 0;
}

int velo_clustering_check(uint32_t n_events, const uint32_t* found_offsets, const uint32_t* found_ids,
                          const uint32_t* expected_offsets, const uint32_t* expected_ids, int just_check_ids,
                          float allowed_distance_error, velo_clustering_check_result* r)
{
  if (!found_offsets || !expected_offsets || !r) return -1;
  std::memset(r, 0, sizeof(*r));
  const bool just_ids = just_check_ids != 0;
  for (uint32_t e = 0; e < n_events; ++e) {
    const uint32_t* f = found_ids + found_offsets[e];
    const uint32_t nf = found_offsets[e + 1] - found_offsets[e];
    const uint32_t* x = expected_ids + expected_offsets[e];
    const uint32_t nx = expected_offsets[e + 1] - expected_offsets[e];
    r->reconstructible += nx;
    r->reconstructed += nf;
    std::set<uint32_t> expected(x, x + nx);
    for (uint32_t i = 0; i < nf; ++i) {
      bool ok = expected.count(f[i]) != 0;
      for (uint32_t k = 0; !ok && !just_ids && k < nx; ++k) ok = matches(f[i], x[k], false, allowed_distance_error);
      if (ok) ++r->correct;
      else ++r->ghosts;
    }
    // clones: per distinct found ID, the found clusters that match it beyond the first
    const std::set<uint32_t> distinct(f, f + nf);
    for (const uint32_t id : distinct) {
      uint64_t reps = 0;
      for (uint32_t i = 0; i < nf; ++i) reps += matches(id, f[i], just_ids, allowed_distance_error) ? 1u : 0u;
      if (reps > 1) r->clones += reps - 1;
    }
  }
  if (r->reconstructible) r->efficiency = (float) r->correct / (float) r->reconstructible;
  if (r->reconstructed) {
    r->clone_fraction = (float) r->clones / (float) r->reconstructed;
    r->ghost_fraction = (float) r->ghosts / (float) r->reconstructed;
  }
  return 0;
}

} // extern "C"
