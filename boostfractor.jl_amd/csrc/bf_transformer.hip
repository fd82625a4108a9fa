// bf_transformer.hip -- mode matching (placeholder, filled in below)
#include "bf_internal.h"
namespace bf {
[[noreturn]] static void todo() { throw Error(BF_ERR_UNSUPPORTED, "not implemented yet"); }
void seed_modes_dev(bf_ctx*, const bf_coords*, int, int, double, double*, double*) { todo(); }
void field2modes_dev(bf_ctx*, const bf_coords*, const bf_modes*, const double*, double, double, double, double*) { todo(); }
void modes2field_dev(bf_ctx*, const bf_coords*, const bf_modes*, const double*, double*) { todo(); }
void propagation_matrices_dev(bf_ctx*, const bf_setup*, const bf_coords*, const bf_modes*, const double*, int, int, double, int, double*) { todo(); }
void transformer_dev(bf_ctx*, const bf_setup*, const bf_coords*, const bf_modes*, const double*, int, int, const double*, double, const double*, const double*, int, double*, double*) { todo(); }
void trace_back_dev(bf_ctx*, const double*, const double*, const bf_setup*, const bf_coords*, const bf_modes*, double, int, const double*, double, int, const double*, double*) { todo(); }
}
