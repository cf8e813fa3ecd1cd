// Mersenne-Twister states at given positions of a seeded stream (host only).
//
// Every random start of the solver is ``torch.randn(size, generator=torch.Generator().manual_seed(s),
// dtype=float64)`` drawn on the CPU and then moved to the device (reference src/cdft/precision.py:
// 246-263), so a GPU run starts from exactly the CPU run's vectors. torch draws those serially at
// about 30 ns a value, ~40 ms of one core for a 5 x 65^3 start block. The draw splits exactly at
// any multiple of 16 values (each group of 16 uniforms is transformed on its own), so the host side
// cuts a draw into chunks and lets torch draw every chunk on its own thread from the generator state
// at the chunk's first uniform. This file supplies those states: MT19937 (Matsumoto & Nishimura
// 1998, the generator torch's CPU generator runs) seeded by init_genrand with the low 32 bits of the
// seed, advanced by whole twists, at a list of positions counted in 32-bit outputs (two per float64
// uniform). Advancing costs one twist per 624 outputs and no tempering or conversion, ~1 ns a word.

#include <cstdint>
#include <cstring>

#include "common.h"

namespace cdft {
namespace {

constexpr int kN = 624;
constexpr int kM = 397;

void init_genrand(uint32_t* mt, uint32_t seed) {
  mt[0] = seed;
  for (int i = 1; i < kN; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
}

void twist(uint32_t* mt) {
  auto mix = [](uint32_t u, uint32_t v) {
    const uint32_t y = (u & 0x80000000u) | (v & 0x7fffffffu);
    return (y >> 1) ^ ((v & 1u) ? 0x9908b0dfu : 0u);
  };
  int k = 0;
  for (; k < kN - kM; ++k) mt[k] = mt[k + kM] ^ mix(mt[k], mt[k + 1]);
  for (; k < kN - 1; ++k) mt[k] = mt[k + kM - kN] ^ mix(mt[k], mt[k + 1]);
  mt[kN - 1] = mt[kM - 1] ^ mix(mt[kN - 1], mt[0]);
}

}  // namespace
}  // namespace cdft

using namespace cdft;

// For each position p[i] (ascending, in 32-bit outputs already consumed from the freshly seeded
// stream): key[624 i .. 624 i + 623] = the state array and index[i] = the index of the next word to
// output (624 = a twist is due before the next output).
extern "C" int cdft_mt19937_states(uint64_t seed, int n, const int64_t* positions, uint32_t* key, int32_t* index) {
  if (n < 0 || (n > 0 && (positions == nullptr || key == nullptr || index == nullptr)))
    return fail(CDFT_E_ARG, "mt19937_states: null argument");
  uint32_t mt[kN];
  init_genrand(mt, static_cast<uint32_t>(seed & 0xffffffffu));
  int64_t twists = 0;  // twists applied to mt
  for (int i = 0; i < n; ++i) {
    const int64_t p = positions[i];
    if (p < 0 || (i > 0 && p < positions[i - 1]))
      return fail(CDFT_E_ARG, "mt19937_states: positions must be ascending and non-negative");
    // after p outputs the generator has twisted ceil(p / 624) times and sits at p - 624 (twists - 1)
    const int64_t need = (p + kN - 1) / kN;
    for (; twists < need; ++twists) twist(mt);
    std::memcpy(key + static_cast<size_t>(kN) * i, mt, sizeof(mt));
    index[i] = p == 0 ? kN : static_cast<int32_t>(p - static_cast<int64_t>(kN) * (need - 1));
  }
  return CDFT_OK;
}
