/* Initial-condition selection for a pseudopotential reference build whose drop
 * crosses both periodic lines: the reference picks its initialiser by editing
 * Ark/Constant/ZeroCondition.h, which main() includes textually
 * (Ark/Cpp/main.cpp:130).  This one runs SCMP_Drop (Ark/Cpp/Preprocess.cpp:337)
 * and then shifts the state cyclically by (60, 70) cells, so that the periodic
 * faces (stored twice, zones 12 and 8) carry non-zero forces.  Only Rho, p and
 * f.Tilde vary from cell to cell after SCMP_Drop.  Test infrastructure only. */
SCMP_Drop();
{
    const int sx = 60, sy = 70, ncell = Nx * Ny;
    double *rho = new double[ncell], *pr = new double[ncell], *ft = new double[ncell * DV_Qv];
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            Cell_2D &c = *CarCellArray[i + 1][j + 1];
            rho[j * Nx + i] = c.MsQ().Rho;
            pr[j * Nx + i] = c.MsQ().p;
            for (int q = 0; q < DV_Qv; ++q) ft[(j * Nx + i) * DV_Qv + q] = c.f.Tilde[0][q];
        }
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            Cell_2D &c = *CarCellArray[(i + sx) % Nx + 1][(j + sy) % Ny + 1];
            c.MsQ().Rho = rho[j * Nx + i];
            c.MsQ().p = pr[j * Nx + i];
            for (int q = 0; q < DV_Qv; ++q) c.f.Tilde[0][q] = ft[(j * Nx + i) * DV_Qv + q];
        }
    delete[] rho;
    delete[] pr;
    delete[] ft;
}
