/*
 * rows_io.cpp — persistence of assembled G/B rows per node batch, and the resumable
 * row assembly built on it.
 *
 * The reference persists only the finished system (Escribe_GB_COO / Escribe_GB_CSR,
 * Mesh/SistemaLineal.cpp:96-138) and its solution (solution_GMRES.txt,
 * Solvers/SistemaLinealSolverGMRES_OMP.hpp:606); a run that stops loses every row.
 * Here a batched assembly writes each node batch's rows to its own file as soon as
 * they exist, and a later run with the same inputs reads the files it finds
 * instead of recomputing those batches (rows_resume.cpp).  Rows depend only on
 * (seed, node stream id, path id), so a resumed assembly equals an uninterrupted
 * one bit for bit.  This file is host code only.
 *
 * File (little-endian):  "SPACROWS" | u32 version | u32 reserved | i64 batch index |
 *   i32 nodes | i32 reserved | u64 seed | u64 input hash | i64 G entries |
 *   u64 payload hash | payload: i64 G_offset[nodes + 1] (from 0) | f64 B[nodes] |
 *   f64 G[entries] | i32 G_col[entries] | spac_point_stats[nodes].
 * Hashes are FNV-1a (64 bit).  A file is written to "<name>.tmp" and renamed, so a
 * reader sees either a whole file or none.
 */
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spac_fk.h"
#include "../csrc/spac_error.h"
#include "rows_io.hpp"

namespace {

constexpr char kMagic[8] = {'S', 'P', 'A', 'C', 'R', 'O', 'W', 'S'};
constexpr uint32_t kVersion = 1;

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t reserved0;
    int64_t batch;
    int32_t nodes;
    int32_t reserved1;
    uint64_t seed;
    uint64_t input_hash;
    int64_t entries;
    uint64_t payload_hash;
};
static_assert(sizeof(FileHeader) == 64, "row file header layout");

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    }
};

size_t payload_bytes(int32_t n, int64_t entries) {
    return (size_t)(n + 1) * 8 + (size_t)n * 8 + (size_t)entries * 12 + (size_t)n * sizeof(spac_point_stats);
}

/* The rows of one batch, laid out as in the file's payload. */
struct Payload {
    std::vector<unsigned char> bytes;
    int64_t* goff = nullptr;
    double* B = nullptr;
    double* G = nullptr;
    int32_t* col = nullptr;
    spac_point_stats* st = nullptr;
    void shape(int32_t n, int64_t entries) {
        bytes.assign(payload_bytes(n, entries), 0);
        unsigned char* p = bytes.data();
        goff = (int64_t*)p; p += (size_t)(n + 1) * 8;
        B = (double*)p; p += (size_t)n * 8;
        G = (double*)p; p += (size_t)entries * 8;
        col = (int32_t*)p; p += (size_t)entries * 4;
        st = (spac_point_stats*)p;
    }
};

spac_status write_file(const std::string& path, int64_t batch, uint64_t seed, uint64_t input_hash, int32_t n,
                       int64_t entries, const unsigned char* payload) {
    FileHeader hd;
    std::memset(&hd, 0, sizeof hd);
    std::memcpy(hd.magic, kMagic, 8);
    hd.version = kVersion;
    hd.batch = batch;
    hd.nodes = n;
    hd.seed = seed;
    hd.input_hash = input_hash;
    hd.entries = entries;
    Fnv f;
    f.add(payload, payload_bytes(n, entries));
    hd.payload_hash = f.h;
    const std::string tmp = path + ".tmp";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return spac::set_error(SPAC_E_INVALID_ARG, "cannot write %s", tmp.c_str());
    bool ok = std::fwrite(&hd, sizeof hd, 1, fp) == 1 &&
              std::fwrite(payload, 1, payload_bytes(n, entries), fp) == payload_bytes(n, entries);
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return spac::set_error(SPAC_E_INVALID_ARG, "cannot write %s", path.c_str());
    }
    return SPAC_OK;
}

/* 1: read and valid; 0: missing; -1: present but not these rows (other inputs, or damaged) */
int read_file(const std::string& path, int64_t batch, uint64_t seed, uint64_t input_hash, int32_t n, int64_t entries,
              Payload* out) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) return 0;
    FileHeader hd;
    bool ok = std::fread(&hd, sizeof hd, 1, fp) == 1 && std::memcmp(hd.magic, kMagic, 8) == 0 &&
              hd.version == kVersion && hd.batch == batch && hd.nodes == n && hd.seed == seed &&
              hd.input_hash == input_hash && hd.entries == entries;
    if (ok) {
        out->shape(n, entries);
        ok = std::fread(out->bytes.data(), 1, out->bytes.size(), fp) == out->bytes.size();
        unsigned char extra;
        ok = ok && std::fread(&extra, 1, 1, fp) == 0; /* nothing after the payload */
        if (ok) {
            Fnv f;
            f.add(out->bytes.data(), out->bytes.size());
            ok = f.h == hd.payload_hash;
        }
    }
    std::fclose(fp);
    return ok ? 1 : -1;
}

} // namespace

uint64_t spac::rows_batch_input_hash(const std::vector<char>& table_sig, const spac_node_batch* nb, int32_t b0, int32_t n,
                          uint64_t seed) {
    Fnv f;
    f.add(table_sig.data(), table_sig.size());
    f.add(&seed, 8);
    f.add(&n, 4);
    f.add(nb->x + b0, (size_t)n * 8);
    f.add(nb->y + b0, (size_t)n * 8);
    f.add(nb->iface + b0, (size_t)n * 4);
    f.add(nb->frontera + b0, (size_t)n * 4);
    f.add(nb->paths + b0, (size_t)n * 4);
    f.add(nb->h + b0, (size_t)n * 8);
    f.add(nb->stream + b0, (size_t)n * 4);
    return f.h;
}

std::string spac::rows_batch_path(const char* dir, int64_t b) {
    char name[64];
    std::snprintf(name, sizeof name, "/rows_%06lld.bin", (long long)b);
    return std::string(dir) + name;
}

extern "C" {

spac_status spac_rows_file_write(const char* path, int64_t batch, uint64_t seed, uint64_t input_hash, int32_t n,
                                 const spac_row_block* rows) {
    if (!path || !rows || n < 1 || !rows->B || !rows->G || !rows->G_col || !rows->G_offset || !rows->stats)
        return spac::set_error(SPAC_E_INVALID_ARG, "bad arguments");
    const int64_t base = rows->G_offset[0], entries = rows->G_offset[n] - base;
    if (entries < 0) return spac::set_error(SPAC_E_INVALID_ARG, "G_offset decreases");
    Payload p;
    p.shape(n, entries);
    for (int k = 0; k <= n; ++k) p.goff[k] = rows->G_offset[k] - base;
    std::memcpy(p.B, rows->B, (size_t)n * 8);
    std::memcpy(p.G, rows->G + base, (size_t)entries * 8);
    std::memcpy(p.col, rows->G_col + base, (size_t)entries * 4);
    std::memcpy(p.st, rows->stats, (size_t)n * sizeof(spac_point_stats));
    return write_file(path, batch, seed, input_hash, n, entries, p.bytes.data());
}

spac_status spac_rows_file_read(const char* path, int64_t batch, uint64_t seed, uint64_t input_hash, int32_t n,
                                int64_t entries, spac_row_block* rows, int32_t* found) {
    if (!path || !rows || !found || n < 1 || entries < 0 || !rows->B || !rows->G || !rows->G_col || !rows->G_offset)
        return spac::set_error(SPAC_E_INVALID_ARG, "bad arguments");
    Payload p;
    *found = read_file(path, batch, seed, input_hash, n, entries, &p);
    if (*found != 1) return SPAC_OK;
    const int64_t base = rows->G_offset[0];
    for (int k = 0; k <= n; ++k) rows->G_offset[k] = base + p.goff[k];
    std::memcpy(rows->B, p.B, (size_t)n * 8);
    std::memcpy(rows->G + base, p.G, (size_t)entries * 8);
    std::memcpy(rows->G_col + base, p.col, (size_t)entries * 4);
    if (rows->stats) std::memcpy(rows->stats, p.st, (size_t)n * sizeof(spac_point_stats));
    return SPAC_OK;
}

} /* extern "C" */
