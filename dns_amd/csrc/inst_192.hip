// Engine instantiation for n = 192 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_192(int prec) { return make_engine<192>(prec); }
}  // namespace dnsk
