/*
 * rows_io.hpp — helpers shared by the row-file format (rows_io.cpp) and the resumable row
 * assembly (rows_resume.cpp).
 */
#ifndef SPAC_ROWS_IO_HPP
#define SPAC_ROWS_IO_HPP

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spac_fk.h"

namespace spac {
/* the interface-table bytes a row assembly reads (fk_rows.hip) */
std::vector<char> table_signature(const spac_interface_table* t);
/* FNV-1a of the table signature, seed and the batch's node arrays [b0, b0 + n) */
uint64_t rows_batch_input_hash(const std::vector<char>& table_sig, const spac_node_batch* nb, int32_t b0, int32_t n,
                               uint64_t seed);
/* dir/rows_<b, six digits>.bin */
std::string rows_batch_path(const char* dir, int64_t b);
} // namespace spac

#endif
