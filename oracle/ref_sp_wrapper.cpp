// Test infrastructure: a C-ABI wrapper around the reference's own
// cache_sp_patterns (x86/velo/clustering/src/ClusteringCommon.cpp:3-101),
// compiled together with that file into oracle/_ref/libref_sp.so so the
// oracle's SP tables can be checked against the reference itself.
#include <cstdint>
#include <cstring>
#include <vector>

void cache_sp_patterns(
  std::vector<unsigned char>& sp_patterns,
  std::vector<unsigned char>& sp_sizes,
  std::vector<float>& sp_fx,
  std::vector<float>& sp_fy);

extern "C" void ref_cache_sp_patterns(uint8_t* patterns, uint8_t* sizes, float* fx, float* fy)
{
  // Constants.cu:146-150 sizes and zero-initialises the tables this way.
  std::vector<unsigned char> p(256, 0), s(256, 0);
  std::vector<float> x(512, 0), y(512, 0);
  cache_sp_patterns(p, s, x, y);
  std::memcpy(patterns, p.data(), 256);
  std::memcpy(sizes, s.data(), 256);
  std::memcpy(fx, x.data(), 512 * sizeof(float));
  std::memcpy(fy, y.data(), 512 * sizeof(float));
}
