"""The time-series record of the reference's time loop (time_dependant.dat,
Source/NS3D_periodic.cu:1292-1452): after every RK3 step, 50 columns of
probe values, diagnostics and Fourier coefficients.

Per step (all on the GPU through the library; the point values are gathered
into one small device tensor and copied to the host once):
  u      = iFFTN(u_hat)                      Image_to_Domain
  u_filt = iFFTN(CutOff(u_hat, 0.1))         the filtered field the loop keeps
                                             (its Helmholtz-filtered field is
                                             overwritten by the cut-off one)
  rot    = get_curl(u), rot_filt = get_curl(u_filt)
  E, D   = get_kinetic_energy(u), get_dissipation(u) (the reference formulas)
Columns (1-based, as the reference's comments number them):
  1        time (the loop adds dt twice per step, so column 1 is 2 dt (t+1);
           ``time_increments=1`` gives the physical time instead)
  2-7      u, rot at P1 = (Nx/2-4, Ny/2-3, Nz/2-2);   8  E
  9-14     u, rot at P2 = (Nx/3, Ny/3, Nz/3);          15 D
  16-21    u, rot at P3 = (Nx/8, Ny/8, Nz/8);          22 E - D/Re
  23-28    u_filt, rot_filt at P2;                     29 E
  30-35    Re, Im of ux, uy, uz hat at (Nx-6, Ny-6, Mz-6); 36 D
  37-42    Im ux, uy, uz hat at F1 = (Nx/5, Ny/4-3, Mz/4-1) and F2 = (Nx/5, Ny/5, Mz/5)
           (ux F1, ux F2, uy F1, uy F2, uz F1, uz F2);  43 Im ux hat at (2,2,2)
  44-50    Im ux hat at (1,0,0), (0,0,0), (1,1,0), (1,1,1), (Nx-1,0,0), (Nx-1,Ny-1,0),
           (Nx-1,Ny-1,Mz-1)
The spectral indices use Mz where the reference sets Nz := Mz before reading
them (:1397-1398).  Integer divisions are C's.
"""
import numpy as np


class TimeSeries:
    """Collects the 50-column rows of time_dependant.dat for one NS3D context."""

    NCOL = 50

    def __init__(self, g, time_increments=2, cutoff_radius=0.1):
        self.g = g
        self.inc = int(time_increments)
        self.radius = cutoff_radius
        self.total_time = 0.0
        self.rows = []
        Nx, Ny, Nz, Mz = g.Nx, g.Ny, g.Nz, g.Mz
        self.P1 = (Nx // 2 - 4, Ny // 2 - 3, Nz // 2 - 2)
        self.P2 = (Nx // 3, Ny // 3, Nz // 3)
        self.P3 = (Nx // 8, Ny // 8, Nz // 8)
        self.S0 = (Nx - 6, Ny - 6, Mz - 6)
        self.F1 = (Nx // 5, Ny // 4 - 3, Mz // 4 - 1)
        self.F2 = (Nx // 5, Ny // 5, Mz // 5)
        self.FL = [(2, 2, 2), (1, 0, 0), (0, 0, 0), (1, 1, 0), (1, 1, 1), (Nx - 1, 0, 0), (Nx - 1, Ny - 1, 0),
                   (Nx - 1, Ny - 1, Mz - 1)]
        for p in (self.P1, self.P2, self.P3, self.S0, self.F1, self.F2, *self.FL):
            if min(p) < 0:
                raise ValueError(f"grid {Nx}x{Ny}x{Nz} too small for the reference's probe points ({p})")

    def record(self, u_hat, dt):
        """Append the row of the state u_hat after a step of length dt; returns it."""
        import torch
        g = self.g
        u = [g.iFFTN(c) for c in u_hat]
        filt = [g.iFFTN(c) for c in g.CutOff_Fourier_Filter(u_hat, self.radius)]
        rot = g.get_curl(u)
        rotf = g.get_curl(filt)
        E = g.get_kinetic_energy(u)
        D = g.get_dissipation(u)
        self.total_time += self.inc * dt

        def at(fields, p):
            return [f[p] for f in fields]

        vals = []
        vals += at(u, self.P1) + at(rot, self.P1)
        vals += at(u, self.P2) + at(rot, self.P2)
        vals += at(u, self.P3) + at(rot, self.P3)
        vals += at(filt, self.P2) + at(rotf, self.P2)
        re = [torch.view_as_real(c) for c in u_hat]      # [..., 2]: real, imaginary
        for c in re:
            vals += [c[self.S0][0], c[self.S0][1]]
        for c in re:
            vals += [c[self.F1][1], c[self.F2][1]]
        vals += [re[0][p][1] for p in self.FL]
        v = torch.stack(vals).cpu().numpy()
        row = np.empty(self.NCOL)
        row[0] = self.total_time
        row[1:7], row[7] = v[0:6], E
        row[8:14], row[14] = v[6:12], D
        row[15:21], row[21] = v[12:18], E - D / g.Re
        row[22:28], row[28] = v[18:24], E
        row[29:35], row[35] = v[24:30], D
        row[36:42], row[42] = v[30:36], v[36]
        row[43:50] = v[37:44]
        self.rows.append(row)
        return row

    def array(self):
        return np.array(self.rows).reshape(-1, self.NCOL)

    def write(self, path):
        """time_dependant.dat: one line per step, "%le" fields separated by spaces."""
        with open(path, "w") as fh:
            for r in self.rows:
                fh.write(" ".join("%e" % x for x in r) + "\n")
