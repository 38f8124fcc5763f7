// Engine instantiation for n = 512 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_512(int prec) { return make_engine<512>(prec); }
}  // namespace dnsk
