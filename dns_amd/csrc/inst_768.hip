// Engine instantiation for n = 768 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_768(int prec) { return make_engine<768>(prec); }
}  // namespace dnsk
