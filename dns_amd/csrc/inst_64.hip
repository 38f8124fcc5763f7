// Engine instantiation for n = 64 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_64(int prec) { return make_engine<64>(prec); }
}  // namespace dnsk
