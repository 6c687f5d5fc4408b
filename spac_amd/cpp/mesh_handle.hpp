/*
 * mesh_handle.hpp — the state behind the opaque spac_mesh handle: a Malla, its
 * row jobs and its interface table flattened for the C ABI.
 */
#ifndef SPAC_MESH_HANDLE_HPP
#define SPAC_MESH_HANDLE_HPP

#include <cstdint>
#include <vector>

#include "../../include/spac_mesh.h"
#include "Malla.hpp"

namespace spac {

struct MeshHandle {
    spac_mesh_desc desc{};
    Malla malla;
    std::vector<TrabajoFila> jobs;
    /* flattened interface table */
    std::vector<double> cx, cy, radius, c2, theta, cond;
    std::vector<int32_t> code, node_offset, node_global;
    std::vector<int64_t> ipsi_offset;
    std::vector<double> ipsi;
    double max_cond = 0.0;
};

} // namespace spac

#endif
