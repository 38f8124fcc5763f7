// Engine instantiation for n = 384 (both precisions); one file per grid size
// so the kernels compile in parallel.
#include "engine.hpp"
namespace dnsk {
Impl *make_engine_384(int prec) { return make_engine<384>(prec); }
}  // namespace dnsk
