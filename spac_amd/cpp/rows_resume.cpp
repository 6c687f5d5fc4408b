/*
 * rows_resume.cpp — spac_fk_assemble_rows_resumable: a row assembly in node batches, each
 * batch's rows read from its file (rows_io.cpp) when a file of these inputs exists, else
 * assembled on the device and written.
 */
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spac_fk.h"
#include "../csrc/spac_error.h"
#include "rows_io.hpp"

extern "C" {

spac_status spac_fk_assemble_rows_resumable(spac_problem* prob, const spac_interface_table* t,
                                            const spac_node_batch* nb, uint64_t seed, int32_t batch_nodes,
                                            const char* dir, spac_row_block* out, spac_resume_report* report) {
    if (!prob || !t || !nb || !dir || !out || !out->B || !out->G || !out->G_col || !out->G_offset)
        return spac::set_error(SPAC_E_INVALID_ARG, "null argument");
    if (batch_nodes < 1) return spac::set_error(SPAC_E_INVALID_ARG, "batch_nodes must be >= 1");
    if (nb->n < 1 || !nb->x || !nb->y || !nb->iface || !nb->frontera || !nb->paths || !nb->h || !nb->stream)
        return spac::set_error(SPAC_E_INVALID_ARG, "node batch is incomplete");
    if (t->n_interfaces < 1 || !t->node_offset) return spac::set_error(SPAC_E_INVALID_ARG, "interface table is incomplete");
    const int32_t n = nb->n;
    /* the output's G offsets: each row holds its interface's node count of weights */
    std::vector<int64_t> goff(n + 1, 0);
    for (int k = 0; k < n; ++k) {
        const int f = nb->iface[k];
        if (f < 0 || f >= t->n_interfaces)
            return spac::set_error(SPAC_E_INVALID_ARG, "node %d: interface %d out of range", k, f);
        goff[k + 1] = goff[k] + (t->node_offset[f + 1] - t->node_offset[f]);
    }
    const std::vector<char> sig = spac::table_signature(t);
    spac_resume_report rep;
    std::memset(&rep, 0, sizeof rep);
    std::vector<spac_point_stats> stats(n);
    std::vector<int64_t> loc;
    for (int32_t b0 = 0, bi = 0; b0 < n; b0 += batch_nodes, ++bi) {
        const int32_t m = std::min(batch_nodes, n - b0);
        const int64_t entries = goff[b0 + m] - goff[b0];
        const uint64_t ih = spac::rows_batch_input_hash(sig, nb, b0, m, seed);
        const std::string path = spac::rows_batch_path(dir, bi);
        loc.assign(goff.begin() + b0, goff.begin() + b0 + m + 1);
        spac_row_block sub;
        sub.B = out->B + b0;
        sub.G = out->G;
        sub.G_col = out->G_col;
        sub.G_offset = loc.data();
        sub.stats = stats.data() + b0;
        int32_t found = 0;
        spac_status rc = spac_rows_file_read(path.c_str(), bi, seed, ih, m, entries, &sub, &found);
        if (rc) return rc;
        rep.batches++;
        if (found == 1) {
            rep.loaded++;
        } else {
            if (found < 0) rep.rejected++;
            /* the batch's nodes keep their stream ids, so its rows are those of the whole batch */
            spac_node_batch part = *nb;
            part.n = m;
            part.x = nb->x + b0;
            part.y = nb->y + b0;
            part.iface = nb->iface + b0;
            part.frontera = nb->frontera + b0;
            part.paths = nb->paths + b0;
            part.h = nb->h + b0;
            part.stream = nb->stream + b0;
            spac_row_block dst;
            dst.B = out->B + b0;
            dst.G = out->G + goff[b0];
            dst.G_col = out->G_col + goff[b0];
            std::vector<int64_t> lo(m + 1);
            dst.G_offset = lo.data();
            dst.stats = stats.data() + b0;
            if ((rc = spac_fk_assemble_rows(prob, t, &part, seed, &dst))) return rc;
            if (lo[m] != entries) return spac::set_error(SPAC_E_INVALID_ARG, "batch %d: %lld G entries, expected %lld",
                                                         bi, (long long)lo[m], (long long)entries);
            spac_row_block wr = sub; /* the batch's rows, offsets from goff[b0] in out's arrays */
            if ((rc = spac_rows_file_write(path.c_str(), bi, seed, ih, m, &wr))) return rc;
            rep.computed++;
        }
    }
    for (int k = 0; k <= n; ++k) out->G_offset[k] = goff[k];
    if (out->stats) std::memcpy(out->stats, stats.data(), (size_t)n * sizeof(spac_point_stats));
    if (report) *report = rep;
    return SPAC_OK;
}

} /* extern "C" */
