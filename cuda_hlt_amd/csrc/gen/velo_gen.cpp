// Seeded synthetic VELO event generator; see include/velo_gen.h.
//
// Each event draws from its own RNG stream (splitmix64 of seed and event
// index), so events are generated in parallel and the output is identical
// for any thread count.
#include "velo_gen.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace {

constexpr int kModules = 52;
constexpr int kSensors = 208;
constexpr int kCols = 768;
constexpr int kRows = 256;
constexpr double kRowPitch = 0.055;
constexpr double kRMin = 5.1;
constexpr double kRMax = 42.0;

// Module z positions, Constants.cu:27-31 (the reference's VELO layout).
constexpr std::array<float, kModules> kModuleZ = {
  -287.5f, -275.f, -262.5f, -250.f, -237.5f, -225.f, -212.5f, -200.f, -137.5f, -125.f, -62.5f, -50.f, -37.5f,
  -25.f,   -12.5f, 0.f,     12.5f,  25.f,    37.5f,  50.f,    62.5f,  75.f,    87.5f,  100.f,  112.5f, 125.f,
  137.5f,  150.f,  162.5f,  175.f,  187.5f,  200.f,  212.5f,  225.f,  237.5f,  250.f,  262.5f, 275.f,  312.5f,
  325.f,   387.5f, 400.f,   487.5f, 500.f,   587.5f, 600.f,   637.5f, 650.f,   687.5f, 700.f,  737.5f, 750.f};

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x)
  {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  Rng(uint64_t seed, uint64_t stream)
  {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next()
  { // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal()
  {
    double u1 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  uint32_t poisson(double mean)
  {
    if (mean <= 0) return 0;
    if (mean < 30) {
      const double l = std::exp(-mean);
      uint32_t k = 0;
      double p = uniform();
      while (p > l) {
        ++k;
        p *= uniform();
      }
      return k;
    }
    const double v = mean + std::sqrt(mean) * normal();
    return v < 0 ? 0u : (uint32_t) std::lround(v);
  }
};

struct SensorFrame {
  double ux, uy, vx, vy, tx, ty, tz; // global = T + lx*u + ly*v
};

struct Geometry {
  std::array<float, kCols> local_x {}, x_pitch {};
  std::array<double, kCols + 1> col_edge {};
  std::array<SensorFrame, kSensors> frames {};
  double length = 0;

  Geometry()
  {
    // Pixel pitch 55 um with wider pixels at the two chip boundaries.
    double acc = 0;
    for (int c = 0; c < kCols; ++c) {
      const bool edge = c == 255 || c == 256 || c == 511 || c == 512;
      const double p = edge ? 0.0825 : 0.055;
      col_edge[c] = acc;
      local_x[c] = (float) (acc + 0.5 * p);
      x_pitch[c] = (float) p;
      acc += p;
    }
    col_edge[kCols] = acc;
    length = acc;
    const double w = kRows * kRowPitch; // 14.08 mm
    // Right-half module (odd): an L of four sensors around the beam hole.
    const SensorFrame right[4] = {
      {0, 1, -1, 0, kRMin + w, -kRMin, -0.25},          // vertical strip, x in [5.1,19.18]
      {1, 0, 0, 1, -kRMin, -kRMin - w, -0.25},          // horizontal strip below the beam
      {0, 1, -1, 0, kRMin + 2 * w, -kRMin, 0.25},       // second vertical strip
      {1, 0, 0, 1, -kRMin, -kRMin - 2 * w, 0.25}};      // second horizontal strip
    for (int m = 0; m < kModules; ++m) {
      const double sgn = (m & 1) ? 1.0 : -1.0; // even modules: rotated by pi
      for (int s = 0; s < 4; ++s) {
        SensorFrame f = right[s];
        f.ux *= sgn;
        f.uy *= sgn;
        f.vx *= sgn;
        f.vy *= sgn;
        f.tx *= sgn;
        f.ty *= sgn;
        f.tz = kModuleZ[m] + f.tz;
        frames[4 * m + s] = f;
      }
    }
  }

  void write(char* out) const
  {
    float* p = reinterpret_cast<float*>(out);
    std::memcpy(p, local_x.data(), sizeof(float) * kCols);
    std::memcpy(p + kCols, x_pitch.data(), sizeof(float) * kCols);
    float* ltg = p + 2 * kCols;
    for (int s = 0; s < kSensors; ++s) {
      const SensorFrame& f = frames[s];
      float* t = ltg + 16 * s;
      for (int k = 0; k < 16; ++k) t[k] = 0.f;
      t[0] = (float) f.ux;
      t[1] = (float) f.vx;
      t[2] = 0.f;
      t[3] = (float) f.uy;
      t[4] = (float) f.vy;
      t[5] = 0.f;
      t[6] = 0.f;
      t[7] = 0.f;
      t[8] = 1.f;
      t[9] = (float) f.tx;
      t[10] = (float) f.ty;
      t[11] = (float) f.tz;
    }
  }

  int column_of(double lx) const
  {
    if (lx < 0 || lx >= length) return -1;
    const auto it = std::upper_bound(col_edge.begin(), col_edge.end(), lx);
    return (int) (it - col_edge.begin()) - 1;
  }
};

const Geometry& geometry()
{
  static const Geometry g;
  return g;
}

struct EventOut {
  std::vector<char> raw;
  std::vector<velo_gen_mcp> mcps;
  std::vector<uint32_t> mc_hits;
};

inline uint32_t lhcb_id(uint32_t sensor, uint32_t cx, uint32_t cy)
{
  return 0x80000000u + ((sensor << 18) | ((cx / 256u) << 16) | ((cx % 256u) << 8) | cy);
}

void generate_event(const velo_gen_config& cfg, uint32_t event, EventOut& out)
{
  const Geometry& g = geometry();
  Rng rng(cfg.seed, event + 1);
  const double nu = cfg.nu_min >= 0 ? cfg.nu_min + (cfg.nu_max - cfg.nu_min) * rng.uniform() : cfg.nu;
  const uint32_t n_pv = rng.poisson(nu);

  // pixel key: sensor << 20 | col << 8 | row
  std::vector<uint32_t> pixels;
  pixels.reserve(8192);
  uint32_t key = 0;
  for (uint32_t v = 0; v < n_pv; ++v) {
    const double pvx = 0.04 * rng.normal();
    const double pvy = 0.04 * rng.normal();
    const double pvz = 45.0 * rng.normal();
    const uint32_t n_tracks = rng.poisson(cfg.tracks_per_pv);
    for (uint32_t t = 0; t < n_tracks; ++t) {
      const double eta = cfg.eta_min + (cfg.eta_max - cfg.eta_min) * rng.uniform();
      const double phi = 2.0 * M_PI * rng.uniform();
      const bool backward = rng.uniform() < cfg.backward_fraction;
      const double theta = 2.0 * std::atan(std::exp(-eta));
      const double tan_t = std::tan(theta);
      const double dirz = backward ? -1.0 : 1.0;
      const double tx = tan_t * std::cos(phi);
      const double ty = tan_t * std::sin(phi);
      const double p = 500.0 + 8000.0 * -std::log(std::max(1e-12, rng.uniform()));
      velo_gen_mcp mcp {};
      mcp.event = event;
      mcp.key = key++;
      mcp.pid = rng.uniform() < 0.5 ? 211 : -211;
      mcp.p = (float) p;
      mcp.pt = (float) (p * std::sin(theta));
      mcp.eta = (float) (backward ? -eta : eta);
      mcp.phi = (float) (phi > M_PI ? phi - 2 * M_PI : phi);
      mcp.n_pv = n_pv;
      mcp.hit_offset = (uint32_t) out.mc_hits.size();
      uint32_t n_hits = 0;
      for (int s = 0; s < kSensors; ++s) {
        const SensorFrame& f = g.frames[s];
        const double dz = f.tz - pvz;
        if (dz * dirz <= 0) continue;
        const double x = pvx + tx * dz * dirz;
        const double y = pvy + ty * dz * dirz;
        const double r = std::sqrt(x * x + y * y);
        if (r < kRMin || r > kRMax) continue;
        const double lx = (x - f.tx) * f.ux + (y - f.ty) * f.uy;
        const double ly = (x - f.tx) * f.vx + (y - f.ty) * f.vy;
        const int col = g.column_of(lx);
        if (col < 0 || ly < 0 || ly >= kRows * kRowPitch) continue;
        const int row = (int) (ly / kRowPitch);
        if (rng.uniform() >= cfg.hit_efficiency) continue;
        // 1-3 pixel cluster: neighbours towards the nearer pixel edge
        const double fxp = (lx - g.col_edge[col]) / g.x_pitch[col];
        const double fyp = ly / kRowPitch - row;
        const double u = rng.uniform();
        const int size = u < 0.5 ? 1 : (u < 0.85 ? 2 : 3);
        int cols[3] = {col, col, col}, rows[3] = {row, row, row};
        const int dcol = fxp < 0.5 ? -1 : 1;
        const int drow = fyp < 0.5 ? -1 : 1;
        const bool along_col = std::fabs(fxp - 0.5) > std::fabs(fyp - 0.5);
        int npx = 1;
        if (size >= 2) {
          if (along_col) cols[npx] = col + dcol;
          else rows[npx] = row + drow;
          ++npx;
        }
        if (size == 3) {
          if (along_col) rows[npx] = row + drow;
          else cols[npx] = col + dcol;
          ++npx;
        }
        int sx = 0, sy = 0, n = 0;
        for (int k = 0; k < npx; ++k) {
          if (cols[k] < 0 || cols[k] >= kCols || rows[k] < 0 || rows[k] >= kRows) continue;
          pixels.push_back(((uint32_t) s << 20) | ((uint32_t) cols[k] << 8) | (uint32_t) rows[k]);
          sx += cols[k];
          sy += rows[k];
          ++n;
        }
        out.mc_hits.push_back(lhcb_id((uint32_t) s, (uint32_t) (sx / n), (uint32_t) (sy / n)));
        ++n_hits;
      }
      mcp.n_hits = n_hits;
      if (n_hits >= 3) mcp.flags |= VELO_MCP_HAS_VELO;
      if (backward) mcp.flags |= VELO_MCP_BACKWARD;
      out.mcps.push_back(mcp);
    }
  }

  // noise pixels, uniform over the active area
  const uint32_t n_noise = rng.poisson(cfg.noise_fraction * (double) pixels.size());
  for (uint32_t k = 0; k < n_noise;) {
    const int s = (int) (rng.uniform() * kSensors) % kSensors;
    const int col = (int) (rng.uniform() * kCols) % kCols;
    const int row = (int) (rng.uniform() * kRows) % kRows;
    const SensorFrame& f = g.frames[s];
    const double lx = g.local_x[col];
    const double ly = (row + 0.5) * kRowPitch;
    const double x = f.tx + lx * f.ux + ly * f.vx;
    const double y = f.ty + lx * f.uy + ly * f.vy;
    const double r = std::sqrt(x * x + y * y);
    if (r < kRMin || r > kRMax) continue;
    pixels.push_back(((uint32_t) s << 20) | ((uint32_t) col << 8) | (uint32_t) row);
    ++k;
  }

  // pixels -> super pixels: (sensor, sp_addr) with sp_addr = sp_col << 6 | sp_row
  std::vector<uint64_t> sps;
  sps.reserve(pixels.size());
  for (uint32_t px : pixels) {
    const uint32_t s = px >> 20, col = (px >> 8) & 0xFFFu, row = px & 0xFFu;
    const uint32_t addr = ((col / 2u) << 6) | (row / 4u);
    const uint32_t bit = (col % 2u) * 4u + (row % 4u);
    sps.push_back(((uint64_t) s << 40) | ((uint64_t) addr << 8) | (1ull << bit));
  }
  std::sort(sps.begin(), sps.end());
  std::vector<uint64_t> merged; // sensor<<40 | addr<<8 | pattern
  for (uint64_t v : sps) {
    if (!merged.empty() && (merged.back() >> 8) == (v >> 8)) merged.back() |= v & 0xFFu;
    else merged.push_back(v);
  }

  // raw event: banks in sensor order, SPs ascending in address
  std::vector<uint32_t> bank_begin(kSensors + 1, 0);
  for (uint64_t v : merged) bank_begin[(v >> 40) + 1]++;
  for (int s = 0; s < kSensors; ++s) bank_begin[s + 1] += bank_begin[s];
  const size_t n_sp = merged.size();
  const size_t size = 4 + 4 * (kSensors + 1) + 8 * kSensors + 4 * n_sp;
  out.raw.assign(size, 0);
  uint32_t* w = reinterpret_cast<uint32_t*>(out.raw.data());
  w[0] = kSensors;
  uint32_t* offs = w + 1;
  uint32_t* payload = w + 1 + kSensors + 1;
  uint32_t pos = 0;
  for (int s = 0; s < kSensors; ++s) {
    offs[s] = pos * 4;
    const uint32_t b = bank_begin[s], e = bank_begin[s + 1];
    payload[pos++] = (uint32_t) s;
    payload[pos++] = e - b;
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t addr = (uint32_t) ((merged[i] >> 8) & 0x7FFFu);
      const uint32_t col = addr >> 6, row = addr & 0x3Fu;
      bool neighbour = false;
      for (int dc = -1; dc <= 1 && !neighbour; ++dc) {
        for (int dr = -1; dr <= 1 && !neighbour; ++dr) {
          if (!dc && !dr) continue;
          const int nc = (int) col + dc, nr = (int) row + dr;
          if (nc < 0 || nr < 0 || nc >= 384 || nr >= 64) continue;
          const uint64_t probe = ((uint64_t) s << 40) | ((uint64_t) (((uint32_t) nc << 6) | (uint32_t) nr) << 8);
          const auto it = std::lower_bound(merged.begin() + b, merged.begin() + e, probe);
          if (it != merged.begin() + e && ((*it) >> 8) == (probe >> 8)) neighbour = true;
        }
      }
      payload[pos++] = (neighbour ? 0u : 0x80000000u) | (addr << 8) | (uint32_t) (merged[i] & 0xFFu);
    }
  }
  offs[kSensors] = pos * 4;
}

struct Handle {
  uint32_t n_events = 0;
  std::vector<char> events;
  std::vector<uint32_t> offsets;
  std::vector<velo_gen_mcp> mcps;
  std::vector<uint32_t> mc_hits;
};

} // namespace

extern "C" {

size_t velo_gen_struct_size(const char* name)
{
  if (!name) return 0;
  if (std::strcmp(name, "velo_gen_config") == 0) return sizeof(velo_gen_config);
  if (std::strcmp(name, "velo_gen_mcp") == 0) return sizeof(velo_gen_mcp);
  return 0;
}

void velo_gen_default_config(velo_gen_config* cfg)
{
  cfg->seed = 0;
  cfg->n_events = 1000;
  cfg->n_threads = 0;
  cfg->nu = 7.6;
  cfg->nu_min = -1.0;
  cfg->nu_max = -1.0;
  cfg->tracks_per_pv = 25.0;
  cfg->eta_min = 2.0;
  cfg->eta_max = 5.0;
  cfg->backward_fraction = 0.2;
  cfg->hit_efficiency = 0.98;
  cfg->noise_fraction = 0.001;
}

void velo_gen_geometry(char* out, size_t size)
{
  if (size < VELO_GEN_GEOMETRY_BYTES) return;
  geometry().write(out);
}

void velo_gen_module_zs(float* out)
{
  for (int m = 0; m < kModules; ++m) out[m] = kModuleZ[m];
}

void* velo_gen_run(const velo_gen_config* cfg)
{
  if (!cfg) return nullptr;
  auto h = std::make_unique<Handle>();
  const uint32_t n = cfg->n_events;
  h->n_events = n;
  std::vector<EventOut> outs(n);
  unsigned nt = cfg->n_threads ? cfg->n_threads : std::max(1u, std::thread::hardware_concurrency());
  nt = std::min<unsigned>(nt, std::max<uint32_t>(1u, n));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    pool.emplace_back([&, t] {
      for (uint32_t e = t; e < n; e += nt) generate_event(*cfg, e, outs[e]);
    });
  }
  for (auto& th : pool) th.join();
  size_t total = 0;
  for (auto& o : outs) total += o.raw.size();
  h->events.resize(total);
  h->offsets.resize(n + 1);
  size_t pos = 0;
  for (uint32_t e = 0; e < n; ++e) {
    h->offsets[e] = (uint32_t) pos;
    std::memcpy(h->events.data() + pos, outs[e].raw.data(), outs[e].raw.size());
    pos += outs[e].raw.size();
    const uint32_t base = (uint32_t) h->mc_hits.size();
    for (auto m : outs[e].mcps) {
      m.hit_offset += base;
      h->mcps.push_back(m);
    }
    h->mc_hits.insert(h->mc_hits.end(), outs[e].mc_hits.begin(), outs[e].mc_hits.end());
    std::vector<char>().swap(outs[e].raw);
  }
  h->offsets[n] = (uint32_t) pos;
  return h.release();
}

uint32_t velo_gen_n_events(const void* h) { return static_cast<const Handle*>(h)->n_events; }
size_t velo_gen_events_size(const void* h) { return static_cast<const Handle*>(h)->events.size(); }
const char* velo_gen_events(const void* h) { return static_cast<const Handle*>(h)->events.data(); }
const uint32_t* velo_gen_offsets(const void* h) { return static_cast<const Handle*>(h)->offsets.data(); }
uint32_t velo_gen_n_mcps(const void* h) { return (uint32_t) static_cast<const Handle*>(h)->mcps.size(); }
const velo_gen_mcp* velo_gen_mcps(const void* h) { return static_cast<const Handle*>(h)->mcps.data(); }
uint32_t velo_gen_n_mc_hits(const void* h) { return (uint32_t) static_cast<const Handle*>(h)->mc_hits.size(); }
const uint32_t* velo_gen_mc_hits(const void* h) { return static_cast<const Handle*>(h)->mc_hits.data(); }
void velo_gen_free(void* h) { delete static_cast<Handle*>(h); }

} // extern "C"
