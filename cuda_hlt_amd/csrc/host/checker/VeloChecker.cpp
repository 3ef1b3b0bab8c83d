// VELO track checker (host).  Association and report semantics follow
// checker/tracking/src/MCAssociator.cpp:13-60, TrackChecker.cpp:93-146 and
// :349-405; the category list is CategoriesTrackChecker.cpp:5-230.
#include "velo_checker.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_set>
#include <utility>
#include <vector>

namespace {

// A category predicate: electron or not, the detector requirement, an
// optional origin flag, an eta selection and a momentum selection.
enum Detector : uint8_t { kVelo, kLong };
enum Eta : uint8_t { kAnyEta, kEta25, kBackward, kForward };
enum Mom : uint8_t { kAnyP, kPBelow5, kPAbove5, kP3Pt400 };

struct Category {
  const char* name;
  bool electron;
  Detector det;
  uint32_t origin; // 0 or one of VELO_MCP_FROM_*
  Eta eta;
  Mom mom;
};

// Same order and labels as TrackCheckerVelo::SetCategories.
constexpr Category kCategories[VELO_CHECKER_N_CATEGORIES] = {
  {"Electrons long eta25", true, kLong, 0, kEta25, kAnyP},
  {"Electrons long fromB eta25", true, kLong, VELO_MCP_FROM_B, kEta25, kAnyP},
  {"Electrons long fromB eta25 p<5GeV", true, kLong, VELO_MCP_FROM_B, kEta25, kPBelow5},
  {"Electrons long fromB eta25 p>3GeV pt>400MeV", true, kLong, VELO_MCP_FROM_B, kEta25, kP3Pt400},
  {"Electrons long fromB eta25 p>5GeV", true, kLong, VELO_MCP_FROM_B, kEta25, kPAbove5},
  {"Electrons long fromD eta25", true, kLong, VELO_MCP_FROM_D, kEta25, kAnyP},
  {"Electrons long fromD eta25 p<5GeV", true, kLong, VELO_MCP_FROM_D, kEta25, kPBelow5},
  {"Electrons long fromD eta25 p>3GeV pt>400MeV", true, kLong, VELO_MCP_FROM_D, kEta25, kP3Pt400},
  {"Electrons long fromD eta25 p>5GeV", true, kLong, VELO_MCP_FROM_D, kEta25, kPAbove5},
  {"Electrons long eta25 p<5GeV", true, kLong, 0, kEta25, kPBelow5},
  {"Electrons long eta25 p>3GeV pt>400MeV", true, kLong, 0, kEta25, kP3Pt400},
  {"Electrons long eta25 p>5GeV", true, kLong, 0, kEta25, kPAbove5},
  {"Electrons long strange eta25", true, kLong, VELO_MCP_FROM_S, kEta25, kAnyP},
  {"Electrons long strange eta25 p<5GeV", true, kLong, VELO_MCP_FROM_S, kEta25, kPBelow5},
  {"Electrons long strange eta25 p>3GeV pt>400MeV", true, kLong, VELO_MCP_FROM_S, kEta25, kP3Pt400},
  {"Electrons long strange eta25 p>5GeV", true, kLong, VELO_MCP_FROM_S, kEta25, kPAbove5},
  {"Electrons Velo", true, kVelo, 0, kAnyEta, kAnyP},
  {"Electrons Velo backward", true, kVelo, 0, kBackward, kAnyP},
  {"Electrons Velo forward", true, kVelo, 0, kForward, kAnyP},
  {"Electrons Velo eta25", true, kVelo, 0, kEta25, kAnyP},
  {"Not electron long eta25", false, kLong, 0, kEta25, kAnyP},
  {"Not electron long fromB eta25", false, kLong, VELO_MCP_FROM_B, kEta25, kAnyP},
  {"Not electron long fromB eta25 p<5GeV", false, kLong, VELO_MCP_FROM_B, kEta25, kPBelow5},
  {"Not electron long fromB eta25 p>3GeV pt>400MeV", false, kLong, VELO_MCP_FROM_B, kEta25, kP3Pt400},
  {"Not electron long fromB eta25 p>5GeV", false, kLong, VELO_MCP_FROM_B, kEta25, kPAbove5},
  {"Not electron long fromD eta25", false, kLong, VELO_MCP_FROM_D, kEta25, kAnyP},
  {"Not electron long fromD eta25 p<5GeV", false, kLong, VELO_MCP_FROM_D, kEta25, kPBelow5},
  {"Not electron long fromD eta25 p>3GeV pt>400MeV", false, kLong, VELO_MCP_FROM_D, kEta25, kP3Pt400},
  {"Not electron long fromD eta25 p>5GeV", false, kLong, VELO_MCP_FROM_D, kEta25, kPAbove5},
  {"Not electron long eta25 p<5GeV", false, kLong, 0, kEta25, kPBelow5},
  {"Not electron long eta25 p>3GeV pt>400MeV", false, kLong, 0, kEta25, kP3Pt400},
  {"Not electron long eta25 p>5GeV", false, kLong, 0, kEta25, kPAbove5},
  {"Not electron long strange eta25", false, kLong, VELO_MCP_FROM_S, kEta25, kAnyP},
  {"Not electron long strange eta25 p<5GeV", false, kLong, VELO_MCP_FROM_S, kEta25, kPBelow5},
  {"Not electron long strange eta25 p>3GeV pt>400MeV", false, kLong, VELO_MCP_FROM_S, kEta25, kP3Pt400},
  {"Not electron long strange eta25 p>5GeV", false, kLong, VELO_MCP_FROM_S, kEta25, kPAbove5},
  {"Not electron Velo", false, kVelo, 0, kAnyEta, kAnyP},
  {"Not electron Velo backward", false, kVelo, 0, kBackward, kAnyP},
  {"Not electron Velo forward", false, kVelo, 0, kForward, kAnyP},
  {"Not electron Velo eta25", false, kVelo, 0, kEta25, kAnyP},
};

bool accepts(const Category& c, const velo_gen_mcp& m)
{
  if (((m.flags & VELO_MCP_IS_ELECTRON) != 0) != c.electron) return false;
  if (c.det == kVelo && !(m.flags & VELO_MCP_HAS_VELO)) return false;
  if (c.det == kLong && !(m.flags & VELO_MCP_IS_LONG)) return false;
  if (c.origin && !(m.flags & c.origin)) return false;
  switch (c.eta) {
  case kEta25:
    if (!(m.eta < 5.f && m.eta > 2.f)) return false;
    break;
  case kBackward:
    if (!(m.eta < 0.f)) return false;
    break;
  case kForward:
    if (!(m.eta > 0.f)) return false;
    break;
  default: break;
  }
  switch (c.mom) {
  case kPBelow5: return m.p < 5e3f;
  case kPAbove5: return m.p > 5e3f;
  case kP3Pt400: return m.p > 3e3f && m.pt > 400.f;
  default: return true;
  }
}

// Running-mean state of one category (TrackChecker.cpp:93-146).
struct Report {
  uint64_t accepted = 0, found = 0, clones = 0, events = 0;
  uint64_t accepted_evt = 0, found_evt = 0;
  double eff_per_event = 0, hit_eff = 0, hit_purity = 0;
  std::unordered_set<size_t> seen;

  void match(size_t key, float weight, size_t n_matched, uint32_t mc_hits)
  {
    if (seen.insert(key).second) {
      ++found;
      ++found_evt;
    }
    else
      ++clones;
    const double n = double(found + clones);
    hit_purity = hit_purity * (n - 1) / n + weight / n;
    const double he = double(n_matched) * weight / double(mc_hits);
    hit_eff = hit_eff * (n - 1) / n + he / n;
  }

  void end_event()
  {
    seen.clear();
    if (accepted_evt) {
      eff_per_event *= double(events) / double(events + 1);
      ++events;
      eff_per_event += (double(found_evt) / double(accepted_evt)) / double(events);
    }
    found_evt = accepted_evt = 0;
  }
};

} // namespace

extern "C" {

const char* velo_checker_category_name(uint32_t i)
{
  return i < VELO_CHECKER_N_CATEGORIES ? kCategories[i].name : nullptr;
}

int velo_checker_run(uint32_t n_events, const uint32_t* event_track_offsets, const uint32_t* track_hit_offsets,
                     const uint32_t* hit_ids, const uint8_t* accepted, const velo_gen_mcp* mcps, size_t n_mcps,
                     const uint32_t* mc_hits, velo_checker_result* result, int32_t* track_mcp, float* track_weight)
{
  if (!event_track_offsets || !track_hit_offsets || !result || (n_mcps && (!mcps || !mc_hits))) return -1;
  if (event_track_offsets[n_events] && !hit_ids) return -1;
  for (size_t i = 1; i < n_mcps; ++i)
    if (mcps[i].event < mcps[i - 1].event) return -1;

  std::vector<Report> reports(VELO_CHECKER_N_CATEGORIES);
  uint64_t tracks = 0, ghosts = 0;
  double ghost_per_event = 0;
  std::vector<std::pair<uint32_t, uint32_t>> id_map; // (LHCbID, particle index)
  std::vector<std::pair<uint32_t, uint32_t>> counts;  // (particle index, hits matched)
  size_t p_lo = 0;

  for (uint32_t e = 0; e < n_events; ++e) {
    size_t p_hi = p_lo;
    while (p_lo < n_mcps && mcps[p_lo].event < e) ++p_lo;
    p_hi = p_lo;
    while (p_hi < n_mcps && mcps[p_hi].event == e) ++p_hi;

    // reconstructible particles per category (TrackEffReport::operator()(mcps))
    for (size_t p = p_lo; p < p_hi; ++p)
      for (uint32_t c = 0; c < VELO_CHECKER_N_CATEGORIES; ++c)
        if (accepts(kCategories[c], mcps[p])) {
          ++reports[c].accepted;
          ++reports[c].accepted_evt;
        }

    // LHCbID -> particle, sorted by LHCbID; the first-listed particle wins
    id_map.clear();
    for (size_t p = p_lo; p < p_hi; ++p)
      for (uint32_t h = 0; h < mcps[p].n_hits; ++h) id_map.emplace_back(mc_hits[mcps[p].hit_offset + h], (uint32_t) p);
    std::stable_sort(id_map.begin(), id_map.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });

    uint64_t tracks_evt = 0, ghosts_evt = 0;
    for (uint32_t t = event_track_offsets[e]; t < event_track_offsets[e + 1]; ++t) {
      if (track_mcp) track_mcp[t] = -1;
      if (track_weight) track_weight[t] = 0.f;
      if (accepted && !accepted[t]) continue;
      ++tracks_evt;
      counts.clear();
      size_t n_matched = 0;
      for (uint32_t h = track_hit_offsets[t]; h < track_hit_offsets[t + 1]; ++h) {
        const uint32_t id = hit_ids[h];
        auto it = std::lower_bound(id_map.begin(), id_map.end(), id,
                                   [](const auto& a, uint32_t v) { return a.first < v; });
        if (it == id_map.end() || it->first != id) continue;
        ++n_matched;
        auto c = std::find_if(counts.begin(), counts.end(), [&](const auto& x) { return x.first == it->second; });
        if (c == counts.end()) counts.emplace_back(it->second, 1u);
        else ++c->second;
      }
      if (counts.empty()) {
        ++ghosts_evt;
        continue;
      }
      // best weight; ties go to the lowest particle index
      auto best = counts.begin();
      for (auto c = counts.begin(); c != counts.end(); ++c)
        if (c->second > best->second || (c->second == best->second && c->first < best->first)) best = c;
      const float weight = float(best->second) / float(n_matched);
      if (track_weight) track_weight[t] = weight;
      if (weight < VELO_CHECKER_MIN_WEIGHT) {
        ++ghosts_evt;
        continue;
      }
      if (track_mcp) track_mcp[t] = (int32_t) best->first;
      const velo_gen_mcp& m = mcps[best->first];
      for (uint32_t c = 0; c < VELO_CHECKER_N_CATEGORIES; ++c)
        if (accepts(kCategories[c], m)) reports[c].match(best->first, weight, n_matched, m.n_hits);
    }
    for (auto& r : reports) r.end_event();
    ghost_per_event *= double(e) / double(e + 1);
    if (tracks_evt) ghost_per_event += (double(ghosts_evt) / double(tracks_evt)) / double(e + 1);
    tracks += tracks_evt;
    ghosts += ghosts_evt;
    p_lo = p_hi;
  }

  std::memset(result, 0, sizeof(*result));
  result->events = n_events;
  result->tracks = tracks;
  result->ghosts = ghosts;
  result->ghost_per_event = ghost_per_event;
  for (uint32_t c = 0; c < VELO_CHECKER_N_CATEGORIES; ++c) {
    auto& o = result->categories[c];
    o.accepted = reports[c].accepted;
    o.found = reports[c].found;
    o.clones = reports[c].clones;
    o.eff_per_event = reports[c].eff_per_event;
    o.hit_eff = reports[c].hit_eff;
    o.hit_purity = reports[c].hit_purity;
  }
  return 0;
}

} // extern "C"
