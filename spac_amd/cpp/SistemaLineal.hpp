/*
 * SistemaLineal.hpp — the sparse interface system G u = B in CSR
 * (Mesh/SistemaLineal.hpp:8-26, Mesh/SistemaLineal.cpp:2-164).
 */
#ifndef SPAC_SISTEMALINEAL_HPP
#define SPAC_SISTEMALINEAL_HPP

#include <string>
#include <vector>

#include "Interfaz.hpp"
#include "Nudo.hpp"

namespace spac {

class SistemaLineal {
public:
    std::vector<std::vector<int>> lista_adyacencia;
    std::vector<double> diagonal;
    int n_filas = 0, n_columnas = 0, nnz = 0;
    std::vector<int> G_i, G_j;
    std::vector<double> G_ij, B, u;

    void Reset_Lista_Adyacencia(int numero_nudos);
    void Anadir_Lista_Adyacencia(const Interfaz& interfaz);
    void Purgar_Lista_Adyacencia();
    void Inicializa_CSR();
    /* Adds nudo.B to B[row] and nudo.G into consecutive slots of the row.
     * Returns 0, or -6 (G column outside the pattern; the reference prints ERROR). */
    int Actualiza_GB(const Nudo& nudo);
    int Actualiza_GB(const Interfaz& interfaz);
    /* Same for a row given as contiguous columns (first_col ... first_col + m - 1). */
    int Actualiza_GB(int fila, double b, const double* g, const int* cols, int m);
    void Convierte_COO(std::vector<int>& G_i_COO, std::vector<int>& G_j_COO, std::vector<double>& G_ij_COO) const;
    /* Text formats of Mesh/SistemaLineal.cpp:96-164 */
    bool Escribe_GB_COO(const std::string& G_path = "Output/G.txt", const std::string& B_path = "Output/B.txt") const;
    bool Escribe_GB_CSR(const std::string& dir = "Output") const;
    /* u, one value per line with 12 decimals (Output/solution_GMRES.txt of the GMRES driver,
     * Solvers/SistemaLinealSolverGMRES_OMP.hpp:606-611) */
    bool Escribe_Solucion(const std::string& path = "Output/solution_GMRES.txt") const;
    bool Lee_GB_COO(const std::string& path = "Output/G_in.txt");
};

} // namespace spac

#endif
