// Reverse-step instantiations for one lattice (kept in its own translation unit so the
// lattices compile in parallel).
#include "adjoint.hpp"

namespace xlbh {
int adjoint_dispatch_d3q27(int store, int compute, const AdjParams& p, hipStream_t s, int mode, uint64_t ws_words,
                          uint8_t* flag, void* lam) {
  return adjoint_dispatch_lat<XLBH_D3Q27>(store, compute, p, s, mode, ws_words, flag, lam);
}
}  // namespace xlbh
