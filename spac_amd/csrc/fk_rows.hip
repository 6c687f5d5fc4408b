/*
 * fk_rows.hip — interface-row (G, B) assembly on the GPU.
 */
#include <hip/hip_runtime.h>

#include "../../include/spac_fk.h"
#include "fk_internal.h"

extern "C" spac_status spac_fk_assemble_rows(spac_problem* hp, const spac_interface_table* ifaces,
                                             const spac_node_batch* nodes, uint64_t seed, spac_row_block* out) {
    (void)hp; (void)ifaces; (void)nodes; (void)seed; (void)out;
    return spac::set_error(SPAC_E_INVALID_ARG, "spac_fk_assemble_rows: not available in this build");
}
