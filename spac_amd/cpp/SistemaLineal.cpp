/*
 * SistemaLineal.cpp — adjacency -> CSR pattern and row accumulation
 * (Mesh/SistemaLineal.cpp:2-95) and the G/B text formats (:96-164).
 */
#include "SistemaLineal.hpp"

#include <algorithm>
#include <cstdio>
#include <set>

namespace spac {

void SistemaLineal::Reset_Lista_Adyacencia(int numero_nudos) {
    lista_adyacencia.assign(numero_nudos, {});
    diagonal.assign(numero_nudos, 0.0);
    n_filas = numero_nudos;
    n_columnas = numero_nudos;
}

void SistemaLineal::Anadir_Lista_Adyacencia(const Interfaz& interfaz) {
    for (const Nudo& in : interfaz.nudos_interior) {
        for (const Nudo& c : interfaz.nudos_circunferencia) {
            lista_adyacencia[in.indice_global].push_back(c.indice_global);
            lista_adyacencia[c.indice_global].push_back(in.indice_global);
        }
        diagonal[in.indice_global] += 1.0;
    }
}

void SistemaLineal::Purgar_Lista_Adyacencia() {
    nnz = 0;
    for (int i = 0; i < (int)lista_adyacencia.size(); ++i) {
        std::set<int> s(lista_adyacencia[i].begin(), lista_adyacencia[i].end());
        s.insert(i);
        lista_adyacencia[i].assign(s.begin(), s.end());
        nnz += (int)lista_adyacencia[i].size();
    }
}

void SistemaLineal::Inicializa_CSR() {
    G_i.assign(n_filas + 1, 0);
    G_j.assign(nnz, 0);
    G_ij.assign(nnz, 0.0);
    B.assign(n_filas, 0.0);
    u.assign(n_filas, 0.0);
    int k = 0;
    for (int fila = 0; fila < n_filas; fila++) {
        G_i[fila] = k;
        for (int col : lista_adyacencia[fila]) {
            G_j[k] = col;
            G_ij[k] = (col == fila) ? diagonal[fila] : 1.0E-23;
            k++;
        }
    }
    G_i[n_filas] = k;
}

int SistemaLineal::Actualiza_GB(int fila, double b, const double* g, const int* cols, int m) {
    B[fila] += b;
    if (m <= 0) return 0;
    int j = G_i[fila];
    for (; j < G_i[fila + 1]; j++)
        if (G_j[j] == cols[0]) break;
    if (j == G_i[fila + 1]) return -6;
    if (j + m > G_i[fila + 1]) return -6;
    for (int q = 0; q < m; ++q) {
        if (G_j[j + q] != cols[q]) return -6; /* columns must be contiguous in the row */
        G_ij[j + q] += g[q];
    }
    return 0;
}

int SistemaLineal::Actualiza_GB(const Nudo& nudo) {
    std::vector<int> cols;
    std::vector<double> vals;
    for (const auto& kv : nudo.G) { cols.push_back(kv.first); vals.push_back(kv.second); }
    return Actualiza_GB(nudo.indice_global, nudo.B, vals.data(), cols.data(), (int)cols.size());
}

int SistemaLineal::Actualiza_GB(const Interfaz& interfaz) {
    int rc = 0;
    for (const Nudo& n : interfaz.nudos_interior) {
        int r = Actualiza_GB(n);
        if (r) rc = r;
    }
    return rc;
}

void SistemaLineal::Convierte_COO(std::vector<int>& Ii, std::vector<int>& Jj, std::vector<double>& V) const {
    Ii.resize(nnz); Jj.resize(nnz); V.resize(nnz);
    for (int fila = 0; fila < n_filas; fila++)
        for (int j = G_i[fila]; j < G_i[fila + 1]; j++) { Ii[j] = fila; Jj[j] = G_j[j]; V[j] = G_ij[j]; }
}

bool SistemaLineal::Escribe_GB_COO(const std::string& G_path, const std::string& B_path) const {
    FILE* fg = std::fopen(G_path.c_str(), "w");
    FILE* fb = std::fopen(B_path.c_str(), "w");
    if (!fg || !fb) { if (fg) std::fclose(fg); if (fb) std::fclose(fb); return false; }
    for (int fila = 0; fila < n_filas; fila++) {
        std::fprintf(fb, "%.8f\n", B[fila]);
        for (int j = G_i[fila]; j < G_i[fila + 1]; j++) std::fprintf(fg, "%d %d %.8f\n", fila, G_j[j], G_ij[j]);
    }
    std::fclose(fg);
    std::fclose(fb);
    return true;
}

bool SistemaLineal::Escribe_GB_CSR(const std::string& dir) const {
    FILE* fi = std::fopen((dir + "/G_i.txt").c_str(), "w");
    FILE* fj = std::fopen((dir + "/G_j.txt").c_str(), "w");
    FILE* fv = std::fopen((dir + "/G_ij.txt").c_str(), "w");
    FILE* fb = std::fopen((dir + "/B.txt").c_str(), "w");
    bool ok = fi && fj && fv && fb;
    if (ok) {
        for (int fila = 0; fila < n_filas; fila++) {
            std::fprintf(fi, "%d\n", G_i[fila]);
            std::fprintf(fb, "%.10f\n", B[fila]);
            for (int j = G_i[fila]; j < G_i[fila + 1]; j++) {
                std::fprintf(fj, "%d\n", G_j[j]);
                std::fprintf(fv, "%.10f\n", G_ij[j]);
            }
        }
    }
    for (FILE* f : {fi, fj, fv, fb}) if (f) std::fclose(f);
    return ok;
}

bool SistemaLineal::Escribe_Solucion(const std::string& path) const {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return false;
    bool ok = true;
    for (size_t i = 0; i < u.size() && ok; ++i) ok = std::fprintf(fp, "%.12f\n", u[i]) > 0;
    return (std::fclose(fp) == 0) && ok;
}

/* Header "nnz n" then "i j v" lines sorted by row (Mesh/SistemaLineal.cpp:139-164).
 * Rows without entries get an empty range (the reference leaves them unset). */
bool SistemaLineal::Lee_GB_COO(const std::string& path) {
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) return false;
    if (std::fscanf(fp, "%d %d", &nnz, &n_filas) != 2) { std::fclose(fp); return false; }
    n_columnas = n_filas;
    G_i.assign(n_filas + 1, 0);
    G_j.assign(nnz, 0);
    G_ij.assign(nnz, 0.0);
    std::vector<int> rows(nnz, 0);
    int ii = 0, it, jt;
    double vt;
    while (ii < nnz && std::fscanf(fp, "%d %d %lf", &it, &jt, &vt) == 3) {
        rows[ii] = it; G_j[ii] = jt; G_ij[ii] = vt; ii++;
    }
    std::fclose(fp);
    if (ii != nnz) return false;
    for (int k = 0; k < nnz; ++k) G_i[rows[k] + 1]++;
    for (int r = 0; r < n_filas; ++r) G_i[r + 1] += G_i[r];
    B.assign(n_filas, 0.0);
    u.assign(n_filas, 0.0);
    return true;
}

} // namespace spac
