// SequenceVisitor::check specialisations of the VELO sequence
// (stream/checkers/include/VeloSequenceCheckers_impl.cuh:8-24): the
// consolidated tracks copied to the host at the end of the run, checked
// against MC truth with the reference's VELO track checker semantics.
#include <stdexcept>
#include <vector>

#include "sequence/SequenceVisitor.h"

template<>
void SequenceVisitor::check<consolidate_velo_tracks_t>(
  const unsigned& start_event_offset,
  const unsigned&,
  const HostBuffers& host_buffers,
  const Constants&,
  CheckerInvoker& invoker) const
{
  if (!invoker.velo_result) return;
  const unsigned n = host_buffers.host_number_of_selected_events[0];
  const unsigned* atomics = host_buffers.host_atomics_velo;
  const unsigned T = atomics[2 * n];
  const unsigned* hit_number = host_buffers.host_velo_track_hit_number;
  const unsigned Ht = hit_number[T];
  // SoA hits: x, y, z, then the LHCbIDs
  const uint32_t* ids = reinterpret_cast<const uint32_t*>(host_buffers.host_velo_track_hits) + 3 * (size_t) Ht;

  // MC particles of each selected event, renumbered to the event's slot
  // (CheckerInvoker.cpp selects MC events through the host event list)
  std::vector<size_t> first(invoker.n_mc_events + 1, invoker.n_mcps);
  for (size_t i = invoker.n_mcps; i-- > 0;)
    if (invoker.mcps[i].event < invoker.n_mc_events) first[invoker.mcps[i].event] = i;
  for (size_t e = invoker.n_mc_events; e-- > 0;)
    if (first[e] > first[e + 1]) first[e] = first[e + 1];
  std::vector<velo_gen_mcp> mcps;
  for (unsigned slot = 0; slot < n; ++slot) {
    const size_t input_event = (size_t) host_buffers.host_event_list[slot] + start_event_offset;
    if (input_event >= invoker.n_mc_events)
      throw std::invalid_argument("check: the run holds more events than the MC truth");
    for (size_t i = first[input_event]; i < invoker.n_mcps && invoker.mcps[i].event == input_event; ++i) {
      velo_gen_mcp m = invoker.mcps[i];
      m.event = slot;
      mcps.push_back(m);
    }
  }
  const int rc = velo_checker_run(
    n, atomics + n, hit_number, ids, host_buffers.host_accepted_velo_tracks, mcps.data(), mcps.size(),
    invoker.mc_hits, invoker.velo_result, nullptr, nullptr);
  if (rc != 0) throw std::invalid_argument("check: the VELO checker rejected the tracks or the MC truth");
  invoker.velo_checked = true;
}
