"""RectilinearGrid and topologies (reference: src/Grids/rectilinear_grid.jl,
src/Grids/areas_volumes_spacings.jl, src/Grids/utils.jl)."""
import numpy as np

from ._abi import PERIODIC, BOUNDED


class AbstractTopology:
    pass


class Periodic(AbstractTopology):
    id = PERIODIC


class Bounded(AbstractTopology):
    id = BOUNDED


def _f32(v):
    return np.float32(v)


class RectilinearGrid:
    """`RectilinearGrid(; size, x, y, z, FT=Float32, topology=(Periodic, Periodic, Bounded),
    landmask=nothing, halo=(2, 2, 2))` — rectilinear_grid.jl:57-122.

    Arrays are 0-based numpy views of the reference's 1-based vectors (dzC[0] is
    the reference's dzC[1]).  Fields are float32 like the reference's default FT.
    """

    def __init__(self, *, size, x, y, z, FT=np.float32, topology=(Periodic, Periodic, Bounded),
                 landmask=None, halo=(2, 2, 2)):
        if FT is not np.float32:
            raise ValueError("only FT = Float32 is implemented on the GPU path")
        Nx, Ny, Nz = (int(s) for s in size)
        Hx, Hy, Hz = (int(h) for h in halo)
        for T in topology:
            if not (isinstance(T, type) and issubclass(T, AbstractTopology)):
                raise ValueError(f"{T} is not a valid topology! Valid topologies are: Periodic and Bounded.")
        self.topology = tuple(topology)
        self.topology_ids = tuple(T.id for T in topology)
        x1, x2 = x
        y1, y2 = y
        assert x1 < x2 and y1 < y2
        Δx = (x2 - x1) / Nx                       # Float64, stored as FT (:72-73)
        Δy = (y2 - y1) / Ny
        self.Nx, self.Ny, self.Nz, self.Hx, self.Hy, self.Hz = Nx, Ny, Nz, Hx, Hy, Hz
        self.Δx, self.Δy = _f32(Δx), _f32(Δy)
        self.xF = np.linspace(-Hx * Δx, (Nx + Hx - 1) * Δx, Nx + 2 * Hx).astype(np.float32)
        self.yF = np.linspace(-Hy * Δy, (Ny + Hy - 1) * Δy, Ny + 2 * Hy).astype(np.float32)
        self.xC = np.linspace((0.5 - Hx) * Δx, (Nx + Hx - 0.5) * Δx, Nx + 2 * Hx).astype(np.float32)
        self.yC = np.linspace((0.5 - Hy) * Δy, (Ny + Hy - 0.5) * Δy, Ny + 2 * Hy).astype(np.float32)
        if isinstance(z, tuple) and len(z) == 2:
            z1, z2 = z
            zz = np.linspace(float(z1), float(z2), Nz + 1).astype(np.float32)
        else:
            zz = np.asarray(z, dtype=np.float64)
            z1, z2 = zz[0], zz[-1]
            zz = zz.astype(np.float32)
        assert z1 > z2, "z must run from the surface downward"
        assert len(zz) == Nz + 1
        P = Nz + 2 * Hz
        zF = np.zeros(P, np.float32)
        zC = np.zeros(P, np.float32)
        dzF = np.zeros(P, np.float32)
        dzC = np.zeros(P, np.float32)
        # rectilinear_grid.jl:98-116 (1-based ranges shifted to 0-based)
        zF[Hz:Nz + Hz + 1] = zz
        zC[Hz:Nz + Hz] = (zF[Hz:Nz + Hz] + zF[Hz + 1:Nz + Hz + 1]) / np.float32(2.0)
        dzF[Hz:Nz + Hz] = zF[Hz:Nz + Hz] - zF[Hz + 1:Nz + Hz + 1]
        dzC[Hz:Nz + Hz - 1] = zC[Hz:Nz + Hz - 1] - zC[Hz + 1:Nz + Hz]
        dzF[0:Hz] = dzF[Hz]
        dzC[0:Hz] = dzF[Hz]
        dzF[Nz + Hz:Nz + 2 * Hz] = dzF[Nz + Hz - 1]
        dzC[Nz + Hz - 1:Nz + 2 * Hz] = dzF[Nz + Hz - 1]
        zF[0:Hz] = zF[Hz:2 * Hz] + dzF[0] * np.float32(Hz)
        zF[Nz + Hz + 1:Nz + 2 * Hz] = zF[Nz + 1:Nz + Hz] - dzF[-1] * np.float32(Hz)
        for i in range(1, Hz + 1):
            zC[Nz + Hz + i - 1] = zC[Nz + Hz + i - 2] - dzF[-1]
        for i in range(Hz, 0, -1):
            zC[i - 1] = zC[i] + dzF[0]
        self.zF, self.zC, self.dzF, self.dzC = zF, zC, dzF, dzC
        self.no_land = landmask is None
        self.landmask = landmask_validation(landmask, Nx, Ny, Nz, Hx, Hy, Hz, self.topology_ids)

    @property
    def padded_shape(self):
        return (self.Nx + 2 * self.Hx, self.Ny + 2 * self.Hy, self.Nz + 2 * self.Hz)

    @property
    def ncells_padded(self):
        a, b, c = self.padded_shape
        return a * b * c

    def interior(self, c):
        """interior(c, grid) — tracer_fields.jl:16."""
        return c[self.Hx:self.Hx + self.Nx, self.Hy:self.Hy + self.Ny, self.Hz:self.Hz + self.Nz]

    # metrics (areas_volumes_spacings.jl:9-22); k is the reference's 1-based padded index
    def ΔxC(self, i, j, k): return self.Δx
    def ΔyC(self, i, j, k): return self.Δy
    def ΔzC(self, i, j, k): return self.dzC[k - 1]
    def ΔxF(self, i, j, k): return self.Δx
    def ΔyF(self, i, j, k): return self.Δy
    def ΔzF(self, i, j, k): return self.dzF[k - 1]
    def Ax(self, i, j, k): return self.Δy * self.dzF[k - 1]
    def Ay(self, i, j, k): return self.Δx * self.dzF[k - 1]
    def Az(self, i, j, k): return self.Δx * self.Δy
    def volume(self, i, j, k): return self.Δx * self.Δy * self.dzF[k - 1]

    def __repr__(self):
        T = tuple(t.__name__ for t in self.topology)
        return (f"RegularRectilinearGrid{{Float32, {T}}}(Nx={self.Nx}, Ny={self.Ny}, Nz={self.Nz}, "
                f"halo=({self.Hx}, {self.Hy}, {self.Hz}))")


def landmask_validation(landmask, Nx, Ny, Nz, Hx, Hy, Hz, topo):
    """Grids/utils.jl:92-129: pad a (Nx,Ny,Nz) Bool landmask with topology-aware halos."""
    if landmask is None:
        landmask = np.ones((Nx, Ny, Nz), dtype=bool)
    else:
        landmask = np.asarray(landmask)
        if landmask.dtype != bool:
            raise ValueError("landmask: type mismatch, typeof(landmask) must be Array{Bool, 3}.")
        if landmask.shape != (Nx, Ny, Nz):
            raise ValueError("landmask: grid mismatch, size(landmask) must equal to (grid.Nx, grid.Ny, grid.Nz).")
    lh = np.zeros((Nx + 2 * Hx, Ny + 2 * Hy, Nz + 2 * Hz), dtype=bool, order="F")
    lh[Hx:Hx + Nx, Hy:Hy + Ny, Hz:Hz + Nz] = landmask
    if topo[0] == PERIODIC:
        lh[0:Hx] = lh[Nx:Nx + Hx].copy()
        lh[Nx + Hx:Nx + 2 * Hx] = lh[Hx:2 * Hx].copy()
    else:
        lh[0:Hx] = lh[Hx:Hx + 1].copy()
        lh[Nx + Hx:Nx + 2 * Hx] = lh[Nx + Hx - 1:Nx + Hx].copy()
    if topo[1] == PERIODIC:
        lh[:, 0:Hy] = lh[:, Ny:Ny + Hy].copy()
        lh[:, Ny + Hy:Ny + 2 * Hy] = lh[:, Hy:2 * Hy].copy()
    else:
        lh[:, 0:Hy] = lh[:, Hy:Hy + 1].copy()
        lh[:, Ny + Hy:Ny + 2 * Hy] = lh[:, Ny + Hy - 1:Ny + Hy].copy()
    lh[:, :, 0:Hz] = lh[:, :, Hz:Hz + 1].copy()
    lh[:, :, Nz + Hz:Nz + 2 * Hz] = lh[:, :, Nz + Hz - 1:Nz + Hz].copy()
    return lh
