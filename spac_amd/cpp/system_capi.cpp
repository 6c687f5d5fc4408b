/*
 * system_capi.cpp — the device-resident interface system of include/spac_mesh.h
 * (spac_mesh_system_*): a mesh's CSR pattern, initial values, job slots and
 * interface table held on the device, and its coarse space for two-level GMRES.
 */
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/spac_mesh.h"
#include "fk_internal.h"
#include "mesh_handle.hpp"
#include "spac_error.h"

using namespace spac;

extern "C" {

#define SYS_HIP_CHECK(expr)                                                                     \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            d->release();                                                                       \
            delete d;                                                                           \
            return set_error(SPAC_E_HIP, "%s failed: %s", #expr, hipGetErrorString(_e));        \
        }                                                                                       \
    } while (0)

spac_status spac_mesh_system_create(const spac_mesh* hm, spac_system** out) {
    if (!hm || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(SPAC_E_NO_DEVICE, "no HIP device visible");
    const MeshHandle* h = reinterpret_cast<const MeshHandle*>(hm);
    const SistemaLineal& s = h->malla.sistema;
    const int n = s.n_filas;
    const int64_t nnz = s.G_i[n];
    /* the values right after Inicializa_CSR (SistemaLineal.cpp: diagonal = containing circles) */
    std::vector<double> init((size_t)nnz);
    for (int r = 0; r < n; ++r)
        for (int j = s.G_i[r]; j < s.G_i[r + 1]; ++j) init[j] = (s.G_j[j] == r) ? s.diagonal[r] : 1.0E-23;
    /* each job's first slot, with Actualiza_GB's checks, and at most one job per slot */
    const int nj = (int)h->jobs.size();
    std::vector<int64_t> slot(nj, -1);
    std::vector<int32_t> job_row(nj), job_iface(nj), job_m(nj), job_front(nj);
    std::vector<uint8_t> hit((size_t)nnz, 0);
    for (int k = 0; k < nj; ++k) {
        const TrabajoFila& jb = h->jobs[k];
        const int r = jb.nudo_global;
        const int o = h->node_offset[jb.interfaz], m = h->node_offset[jb.interfaz + 1] - o;
        job_row[k] = r;
        job_iface[k] = jb.interfaz;
        job_m[k] = m;
        job_front[k] = jb.frontera ? 1 : 0;
        if (jb.frontera) continue;
        if (m <= 0) continue;
        const int32_t* cols = h->node_global.data() + o;
        int j = s.G_i[r];
        for (; j < s.G_i[r + 1]; ++j)
            if (s.G_j[j] == cols[0]) break;
        if (j == s.G_i[r + 1] || j + m > s.G_i[r + 1])
            return set_error(SPAC_E_PATTERN, "row %d: G columns outside the CSR pattern", r);
        for (int q = 0; q < m; ++q) {
            if (s.G_j[j + q] != cols[q]) return set_error(SPAC_E_PATTERN, "row %d: G columns not contiguous", r);
            if (hit[j + q]++) return set_error(SPAC_E_PATTERN, "row %d: two jobs add into CSR slot %d", r, j + q);
        }
        slot[k] = j;
    }
    /* the jobs of every row, in job order (B sums them in that order) */
    std::vector<int32_t> rptr(n + 1, 0), rjobs(nj), cur;
    for (int k = 0; k < nj; ++k) rptr[job_row[k] + 1]++;
    for (int r = 0; r < n; ++r) rptr[r + 1] += rptr[r];
    cur.assign(rptr.begin(), rptr.end() - 1);
    for (int k = 0; k < nj; ++k) rjobs[cur[job_row[k]]++] = k;

    DeviceSystem* d = new DeviceSystem();
    d->n_rows = n;
    d->n_jobs = nj;
    d->nnz = nnz;
    SYS_HIP_CHECK(hipGetDevice(&d->device));
    spac_status rc;
    if ((rc = d->Ip.ensure((size_t)(n + 1) * 4)) || (rc = d->J.ensure((size_t)nnz * 4 + 4)) ||
        (rc = d->init.ensure((size_t)nnz * 8 + 8)) || (rc = d->vals.ensure((size_t)nnz * 8 + 8)) ||
        (rc = d->B.ensure((size_t)n * 8)) || (rc = d->job_slot.ensure((size_t)nj * 8 + 8)) ||
        (rc = d->row_job_ptr.ensure((size_t)(n + 1) * 4)) || (rc = d->row_jobs.ensure((size_t)nj * 4 + 4))) {
        d->release();
        delete d;
        return rc;
    }
    SYS_HIP_CHECK(hipMemcpy(d->Ip.ptr, s.G_i.data(), (size_t)(n + 1) * 4, hipMemcpyHostToDevice));
    if (nnz > 0) {
        SYS_HIP_CHECK(hipMemcpy(d->J.ptr, s.G_j.data(), (size_t)nnz * 4, hipMemcpyHostToDevice));
        SYS_HIP_CHECK(hipMemcpy(d->init.ptr, init.data(), (size_t)nnz * 8, hipMemcpyHostToDevice));
        SYS_HIP_CHECK(hipMemcpy(d->vals.ptr, init.data(), (size_t)nnz * 8, hipMemcpyHostToDevice));
    }
    SYS_HIP_CHECK(hipMemset(d->B.ptr, 0, (size_t)n * 8));
    if (nj > 0) {
        SYS_HIP_CHECK(hipMemcpy(d->job_slot.ptr, slot.data(), (size_t)nj * 8, hipMemcpyHostToDevice));
        SYS_HIP_CHECK(hipMemcpy(d->row_jobs.ptr, rjobs.data(), (size_t)nj * 4, hipMemcpyHostToDevice));
    }
    SYS_HIP_CHECK(hipMemcpy(d->row_job_ptr.ptr, rptr.data(), (size_t)(n + 1) * 4, hipMemcpyHostToDevice));
    /* each row's diagonal slot (the GMRES Jacobi scaling reads it); a pattern with two diagonal
       entries in a row keeps none and GMRES scans the rows instead */
    {
        std::vector<int32_t> ds(n, -1);
        bool unique = true;
        for (int r = 0; r < n && unique; ++r)
            for (int j = s.G_i[r]; j < s.G_i[r + 1]; ++j)
                if (s.G_j[j] == r) {
                    if (ds[r] >= 0) unique = false;
                    ds[r] = j;
                }
        if (unique) {
            if ((rc = d->dslot.ensure((size_t)n * 4 + 4))) {
                d->release();
                delete d;
                return rc;
            }
            SYS_HIP_CHECK(hipMemcpy(d->dslot.ptr, ds.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        }
    }
    /* the interface table and the job arrays the row kernels read */
    spac_interface_table tab;
    spac_mesh_interfaces(hm, &tab);
    std::vector<int32_t> jarr(2 * (size_t)nj);
    std::vector<int64_t> goff(nj + 1, 0);
    for (int k = 0; k < nj; ++k) {
        jarr[k] = job_iface[k];
        jarr[nj + k] = job_front[k];
        goff[k + 1] = goff[k] + job_m[k];
    }
    if ((rc = iface_table_upload(&tab, nullptr, d->t_ifc, d->t_ifoff, d->t_ifnode, d->t_ifipsi)) ||
        (rc = d->t_jobs.ensure(jarr.size() * 4 + 8)) || (rc = d->t_goff.ensure(goff.size() * 8)) ||
        (rc = d->flag.ensure(64))) {
        d->release();
        delete d;
        return rc;
    }
    if (nj > 0) SYS_HIP_CHECK(hipMemcpy(d->t_jobs.ptr, jarr.data(), jarr.size() * 4, hipMemcpyHostToDevice));
    SYS_HIP_CHECK(hipMemcpy(d->t_goff.ptr, goff.data(), goff.size() * 8, hipMemcpyHostToDevice));
    d->n_ifaces = tab.n_interfaces;
    d->rbf_kind = tab.rbf_kind;
    d->stencil_size = tab.stencil_size;
    d->gauss_eps = tab.gauss_eps;
    d->g_total = goff[nj];
    d->mmax = 1;
    for (int k = 0; k < nj; ++k) d->mmax = std::max(d->mmax, job_m[k]);
    /* the fixed part of the every-job batch's solve table (the node batch spac_mesh_jobs gives) */
    {
        std::vector<double> jx(nj), jy(nj), xy(2 * (size_t)nj), circ(3 * (size_t)nj);
        std::vector<int32_t> jr(nj), ji(nj), jf(nj), live((size_t)nj + 1, 0);
        if ((rc = spac_mesh_jobs(hm, jr.data(), jx.data(), jy.data(), ji.data(), jf.data()))) {
            d->release();
            delete d;
            return rc;
        }
        for (int k = 0; k < nj; ++k) {
            xy[k] = jx[k];
            xy[nj + k] = jy[k];
            const int f = job_iface[k];
            circ[3 * (size_t)k] = tab.radius[f];
            circ[3 * (size_t)k + 1] = tab.cx[f];
            circ[3 * (size_t)k + 2] = tab.cy[f];
            live[k + 1] = live[k] + (job_front[k] ? 0 : 1);
        }
        d->n_live = live[nj];
        if ((rc = d->u_xy.ensure(xy.size() * 8 + 8)) || (rc = d->u_circle.ensure(circ.size() * 8 + 8)) ||
            (rc = d->u_live.ensure(live.size() * 4))) {
            d->release();
            delete d;
            return rc;
        }
        if (nj > 0) {
            SYS_HIP_CHECK(hipMemcpy(d->u_xy.ptr, xy.data(), xy.size() * 8, hipMemcpyHostToDevice));
            SYS_HIP_CHECK(hipMemcpy(d->u_circle.ptr, circ.data(), circ.size() * 8, hipMemcpyHostToDevice));
        }
        SYS_HIP_CHECK(hipMemcpy(d->u_live.ptr, live.data(), live.size() * 4, hipMemcpyHostToDevice));
    }
    /* the rows' nodes (spac_mlp_predict_system evaluates the network there) */
    {
        std::vector<double> nx(n), ny(n), nxy(2 * (size_t)n);
        std::vector<int32_t> nf(n);
        if ((rc = spac_mesh_nodes(hm, nx.data(), ny.data(), nf.data())) || (rc = d->node_xy.ensure(nxy.size() * 8 + 8))) {
            d->release();
            delete d;
            return rc;
        }
        for (int i = 0; i < n; ++i) {
            nxy[i] = nx[i];
            nxy[n + i] = ny[i];
        }
        if (n > 0) SYS_HIP_CHECK(hipMemcpy(d->node_xy.ptr, nxy.data(), nxy.size() * 8, hipMemcpyHostToDevice));
    }
    d->h_job_row = std::move(job_row);
    d->h_job_iface = std::move(job_iface);
    d->h_job_m = std::move(job_m);
    d->h_job_frontera = std::move(job_front);
    d->h_Ip.assign(s.G_i.begin(), s.G_i.end());
    d->h_J.assign(s.G_j.begin(), s.G_j.end());
    d->h_table_sig = table_signature(&tab);
    *out = reinterpret_cast<spac_system*>(d);
    return SPAC_OK;
}

spac_status spac_mesh_system_set_coarse(const spac_mesh* hm, spac_system* hs, int32_t nx, int32_t ny) {
    DeviceSystem* d = reinterpret_cast<DeviceSystem*>(hs);
    if (!hm || !d) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (nx == 0 && ny == 0) {
        d->coarse_k = 0;
        return SPAC_OK;
    }
    if (nx < 1 || ny < 1 || nx * ny > kMaxCoarse)
        return set_error(SPAC_E_INVALID_ARG, "coarse grid %d x %d: 1..%d aggregates", nx, ny, kMaxCoarse);
    spac_mesh_info info;
    spac_status rc;
    if ((rc = spac_mesh_get_info(hm, &info))) return rc;
    const int n = d->n_rows;
    if (info.n_nodes != n) return set_error(SPAC_E_INVALID_ARG, "the mesh has %d nodes, the system %d rows", info.n_nodes, n);
    spac_interface_table tab;
    if ((rc = spac_mesh_interfaces(hm, &tab))) return rc;
    if (table_signature(&tab) != d->h_table_sig)
        return set_error(SPAC_E_INVALID_ARG, "the mesh is not the one the system was created from");
    /* no coarse space until the new one is fully in place: a failure below leaves Jacobi GMRES,
       never kernels over half-replaced aggregate arrays */
    d->coarse_k = 0;
    std::vector<double> x(n), y(n);
    std::vector<int32_t> fr(n);
    if ((rc = spac_mesh_nodes(hm, x.data(), y.data(), fr.data()))) return rc;
    double x0 = x[0], x1 = x[0], y0 = y[0], y1 = y[0];
    for (int i = 0; i < n; ++i) {
        x0 = std::min(x0, x[i]); x1 = std::max(x1, x[i]);
        y0 = std::min(y0, y[i]); y1 = std::max(y1, y[i]);
    }
    /* cell of each node on the nx x ny grid over the nodes' bounding box; empty cells dropped */
    std::vector<int32_t> cell(n), id(nx * ny, -1);
    for (int i = 0; i < n; ++i) {
        int cx = x1 > x0 ? (int)((x[i] - x0) / (x1 - x0) * nx) : 0;
        int cy = y1 > y0 ? (int)((y[i] - y0) / (y1 - y0) * ny) : 0;
        cx = std::min(std::max(cx, 0), nx - 1);
        cy = std::min(std::max(cy, 0), ny - 1);
        cell[i] = cx + nx * cy;
    }
    int k = 0;
    for (int i = 0; i < n; ++i)
        if (id[cell[i]] < 0) id[cell[i]] = 1;
    for (int c = 0; c < nx * ny; ++c) id[c] = id[c] > 0 ? k++ : -1;
    std::vector<int32_t> agg(n), ptr(k + 1, 0), rows(n);
    for (int i = 0; i < n; ++i) ptr[(agg[i] = id[cell[i]]) + 1]++;
    for (int a = 0; a < k; ++a) ptr[a + 1] += ptr[a];
    std::vector<int32_t> cur(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) rows[cur[agg[i]]++] = i;
    /* the aggregates each block of kCoarseRowBlock rows spans (the prolongation forms one coarse
       value per aggregate of its block) */
    const int nblk = (n + kCoarseRowBlock - 1) / kCoarseRowBlock;
    std::vector<int32_t> bptr(nblk + 1, 0), bagg;
    std::vector<uint8_t> rslot(n, 0);
    int blk_max = 0;
    for (int b = 0; b < nblk; ++b) {
        const int first = (int)bagg.size();
        for (int i = b * kCoarseRowBlock; i < std::min(n, (b + 1) * kCoarseRowBlock); ++i) {
            int q = first;
            while (q < (int)bagg.size() && bagg[q] != agg[i]) ++q;
            if (q == (int)bagg.size()) bagg.push_back(agg[i]);
            rslot[i] = (uint8_t)std::min(q - first, 255);
        }
        bptr[b + 1] = (int32_t)bagg.size();
        blk_max = std::max(blk_max, bptr[b + 1] - bptr[b]);
    }
    /* each aggregate's entries of the block lists, in block order */
    std::vector<int32_t> aptr(k + 1, 0), aent(bagg.size());
    for (size_t e = 0; e < bagg.size(); ++e) aptr[bagg[e] + 1]++;
    for (int a = 0; a < k; ++a) aptr[a + 1] += aptr[a];
    {
        std::vector<int32_t> fill(aptr.begin(), aptr.end() - 1);
        for (size_t e = 0; e < bagg.size(); ++e) aent[fill[bagg[e]]++] = (int32_t)e;
    }
    if (bagg.empty()) bagg.push_back(0);
    if (aent.empty()) aent.push_back(0);
    if ((rc = d->c_agg.ensure((size_t)n * 4)) || (rc = d->c_ptr.ensure((size_t)(k + 1) * 4)) ||
        (rc = d->c_rows.ensure((size_t)n * 4)) || (rc = d->c_efix.ensure((size_t)k * k * 8)) ||
        (rc = d->c_einv.ensure((size_t)k * k * 8)) || (rc = d->c_work.ensure((size_t)2 * k * k * 8 + 64)) ||
        (rc = d->c_vec.ensure((size_t)2 * k * 8)) || (rc = d->c_bptr.ensure((size_t)(nblk + 1) * 4)) ||
        (rc = d->c_bagg.ensure(bagg.size() * 4)) || (rc = d->c_rslot.ensure((size_t)n + 4)) ||
        (rc = d->c_part.ensure(bagg.size() * 8)) || (rc = d->c_aptr.ensure((size_t)(k + 1) * 4)) ||
        (rc = d->c_aent.ensure(aent.size() * 4)))
        return rc;
    if (hipMemcpy(d->c_agg.ptr, agg.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->c_ptr.ptr, ptr.data(), (size_t)(k + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->c_rows.ptr, rows.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->c_bptr.ptr, bptr.data(), (size_t)(nblk + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->c_bagg.ptr, bagg.data(), bagg.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->c_rslot.ptr, rslot.data(), (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->c_aptr.ptr, aptr.data(), (size_t)(k + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d->c_aent.ptr, aent.data(), aent.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(SPAC_E_HIP, "uploading the coarse space failed");
    d->c_blk_max = blk_max;
    d->coarse_k = k;
    return SPAC_OK;
}

spac_status spac_mesh_system_destroy(spac_system* hs) {
    DeviceSystem* d = reinterpret_cast<DeviceSystem*>(hs);
    if (!d) return SPAC_OK;
    d->release();
    delete d;
    return SPAC_OK;
}

spac_status spac_mesh_system_read(const spac_system* hs, int32_t* G_i, int32_t* G_j, double* G_ij, double* B) {
    const DeviceSystem* d = reinterpret_cast<const DeviceSystem*>(hs);
    if (!d) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (G_i) std::memcpy(G_i, d->h_Ip.data(), d->h_Ip.size() * 4);
    if (G_j) std::memcpy(G_j, d->h_J.data(), d->h_J.size() * 4);
    if (G_ij && d->nnz > 0 && hipMemcpy(G_ij, d->vals.ptr, (size_t)d->nnz * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(SPAC_E_HIP, "reading the system values failed");
    if (B && hipMemcpy(B, d->B.ptr, (size_t)d->n_rows * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(SPAC_E_HIP, "reading the system B failed");
    return SPAC_OK;
}

} /* extern "C" */
