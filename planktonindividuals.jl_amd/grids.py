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


def _deg2rad(v):
    """deg2rad(z::Float64) = z * (π / 180) in binary64."""
    return np.float64(v) * (np.float64(np.pi) / np.float64(180))


def _cos_sin_deg_f32(y, fn):
    """cos / sin of a Float32 angle in degrees as the reference forms it:
    (y * Float32(π)) / 180 in Float32, then the Float32 function (evaluated in
    binary64 and rounded once)."""
    a = (np.asarray(y, np.float32) * np.float32(np.pi)) / np.float32(180)
    return fn(a.astype(np.float64)).astype(np.float32)


class LatLonGrid:
    """`LatLonGrid(; size, lat, lon, z, FT=Float32, radius=6370.0e3, landmask=nothing,
    halo=(2, 2, 2))` — Grids/lat_lon_grid.jl:70-200.  x is Periodic when the
    longitudes span -180..180, else Bounded; y and z are Bounded.  The spacings
    (dxC, dyC, dxF, dyF, Az: (Nx+2Hx, Ny+2Hy)) and the areas and volumes (Ax,
    Ay, Vol: padded 3-D) are Float32 arrays with halos, like the reference's;
    they are handed to the device as they are (pi_metrics)."""

    def __init__(self, *, size, lat, lon, z, FT=np.float32, radius=6370.0e3, landmask=None, halo=(2, 2, 2),
                 _metrics=None):
        if FT is not np.float32:
            raise ValueError("only FT = Float32 is implemented on the GPU path")
        Nx, Ny, Nz = (int(s) for s in size)
        Hx, Hy, Hz = (int(h) for h in halo)
        lat1, lat2 = lat
        lon1, lon2 = lon
        if isinstance(z, tuple) and len(z) == 2:
            z1, z2 = z
            zz = np.linspace(float(z1), float(z2), Nz + 1).astype(np.float32)
        else:
            zz = np.asarray(z, dtype=np.float64)
            z1, z2 = zz[0], zz[-1]
            zz = zz.astype(np.float32)
        if not (-180 <= lon1 < lon2 <= 180):
            raise AssertionError("-180 <= lon₁ < lon₂ <= 180")
        if not (-80 <= lat1 < lat2 <= 80):
            raise AssertionError("-80 <= lat₁ < lat₂ <= 80")
        if not z1 > z2:
            raise AssertionError("z₁ > z₂")
        if len(zz) != Nz + 1:
            raise AssertionError("length(z) == Nz + 1")
        TX = Periodic if (lon1 == -180 and lon2 == 180) else Bounded
        self.topology = (TX, Bounded, Bounded)
        self.topology_ids = tuple(T.id for T in self.topology)
        self.Nx, self.Ny, self.Nz, self.Hx, self.Hy, self.Hz = Nx, Ny, Nz, Hx, Hy, Hz
        Δx = (lon2 - lon1) / Nx                  # degrees, binary64 (:97-98)
        Δy = (lat2 - lat1) / Ny
        self.Δx, self.Δy = _f32(Δx), _f32(Δy)
        Px, Py, P = Nx + 2 * Hx, Ny + 2 * Hy, Nz + 2 * Hz
        self.xF = np.linspace(lon1 - Hx * Δx, lon1 + (Nx + Hx - 1) * Δx, Px).astype(np.float32)
        self.yF = np.linspace(lat1 - Hy * Δy, lat1 + (Ny + Hy - 1) * Δy, Py).astype(np.float32)
        self.xC = np.linspace(lon1 + (0.5 - Hx) * Δx, lon1 + (Nx + Hx - 0.5) * Δx, Px).astype(np.float32)
        self.yC = np.linspace(lat1 + (0.5 - Hy) * Δy, lat1 + (Ny + Hy - 0.5) * Δy, Py).astype(np.float32)
        zF = np.zeros(P, np.float32)
        zC = np.zeros(P, np.float32)
        dzF = np.zeros(P, np.float32)
        dzC = np.zeros(P, np.float32)
        zF[Hz:Nz + Hz + 1] = zz                                          # :120-123
        zC[Hz:Nz + Hz] = (zF[Hz:Nz + Hz] + zF[Hz + 1:Nz + Hz + 1]) / np.float32(2.0)
        dzF[Hz:Nz + Hz] = zF[Hz:Nz + Hz] - zF[Hz + 1:Nz + Hz + 1]
        dzC[Hz:Nz + Hz - 1] = zC[Hz:Nz + Hz - 1] - zC[Hz + 1:Nz + Hz]
        dzF[0:Hz] = dzF[Hz]                                              # :125-130
        dzC[0:Hz] = dzF[Hz]
        dzF[Nz + Hz:] = dzF[Nz + Hz - 1]
        dzC[Nz + Hz - 1:] = dzF[Nz + Hz - 1]
        zF[0:Hz] = zF[Hz:2 * Hz] + dzF[0] * np.float32(Hz)
        zF[Nz + Hz + 1:Nz + 2 * Hz] = zF[Nz + 1:Nz + Hz] - dzF[-1] * np.float32(Hz)
        for i in range(1, Hz + 1):
            zC[Nz + Hz + i - 1] = zC[Nz + Hz + i - 2] - dzF[-1]
        for i in range(Hz, 0, -1):
            zC[i - 1] = zC[i] + dzF[0]
        self.zF, self.zC, self.dzF, self.dzC = zF, zC, dzF, dzC
        if _metrics is not None:                                         # LoadLatLonGrid
            for k, v in _metrics.items():
                setattr(self, k, v)
        else:
            R = np.float64(radius)
            cosC = _cos_sin_deg_f32(self.yC, np.cos)                     # :140-147
            dx_col = (R * cosC.astype(np.float64) * _deg2rad(Δx)).astype(np.float32)
            dy_val = np.float32(R * _deg2rad(Δy))
            self.dxC = np.asfortranarray(np.broadcast_to(dx_col[None, :], (Px, Py)).astype(np.float32))
            self.dxF = self.dxC.copy(order="F")
            self.dyC = np.full((Px, Py), dy_val, np.float32, order="F")
            self.dyF = self.dyC.copy(order="F")
            sinF = _cos_sin_deg_f32(self.yF, np.sin)                     # :148-155
            band = (sinF[1:] - sinF[:-1]).astype(np.float32)             # Float32 difference
            Az = np.zeros((Px, Py), np.float32, order="F")
            Az[:Px - 1, :Py - 1] = (R * R * _deg2rad(Δx) * band.astype(np.float64))[None, :].astype(np.float32)
            Az[Px - 1, :] = Az[Px - 2, :]                                # the corner (Px, Py) stays 0
            Az[:Px - 1, Py - 1] = Az[:Px - 1, Py - 2]
            self.Az = Az
            dz3 = dzF[None, None, :]
            self.Ax = np.asfortranarray((dz3 * self.dyF[:, :, None]).astype(np.float32))   # :157-165
            self.Ay = np.asfortranarray((dz3 * self.dxF[:, :, None]).astype(np.float32))
            self.Vol = np.asfortranarray((self.Az[:, :, None] * dz3).astype(np.float32))
        self.no_land = landmask is None
        self.landmask = landmask_validation(landmask, Nx, Ny, Nz, Hx, Hy, Hz, self.topology_ids)

    padded_shape = RectilinearGrid.padded_shape
    ncells_padded = RectilinearGrid.ncells_padded
    interior = RectilinearGrid.interior

    # metrics (areas_volumes_spacings.jl:24-38); i, j, k are 1-based padded indices
    def ΔxC(self, i, j, k): return self.dxC[i - 1, j - 1]
    def ΔyC(self, i, j, k): return self.dyC[i - 1, j - 1]
    def ΔzC(self, i, j, k): return self.dzC[k - 1]
    def ΔxF(self, i, j, k): return self.dxF[i - 1, j - 1]
    def ΔyF(self, i, j, k): return self.dyF[i - 1, j - 1]
    def ΔzF(self, i, j, k): return self.dzF[k - 1]
    def volume(self, i, j, k): return self.Vol[i - 1, j - 1, k - 1]

    def metric_arrays(self):
        """(dxC, dyC, dxF, dyF, Az, Ax, Ay, Vol) as contiguous column-major Float32 (pi_metrics)."""
        return tuple(np.asfortranarray(a, dtype=np.float32)
                     for a in (self.dxC, self.dyC, self.dxF, self.dyF, self.Az, self.Ax, self.Ay, self.Vol))

    def __repr__(self):
        T = tuple(t.__name__ for t in self.topology)
        return (f"LatLonGrid{{Float32, {T}}}(Nx={self.Nx}, Ny={self.Ny}, Nz={self.Nz}, "
                f"halo=({self.Hx}, {self.Hy}, {self.Hz}))")


# areas_volumes_spacings.jl:9-38 as functions of (i, j, k, grid), 1-based padded indices
def ΔxC(i, j, k, g): return g.ΔxC(i, j, k)
def ΔyC(i, j, k, g): return g.ΔyC(i, j, k)
def ΔzC(i, j, k, g): return g.ΔzC(i, j, k)
def ΔxF(i, j, k, g): return g.ΔxF(i, j, k)
def ΔyF(i, j, k, g): return g.ΔyF(i, j, k)
def ΔzF(i, j, k, g): return g.ΔzF(i, j, k)
def Ax(i, j, k, g): return g.Ax[i - 1, j - 1, k - 1] if isinstance(g, LatLonGrid) else g.Ax(i, j, k)
def Ay(i, j, k, g): return g.Ay[i - 1, j - 1, k - 1] if isinstance(g, LatLonGrid) else g.Ay(i, j, k)
def Az(i, j, k, g): return g.Az[i - 1, j - 1] if isinstance(g, LatLonGrid) else g.Az(i, j, k)
def volume(i, j, k, g): return g.volume(i, j, k)


def LoadLatLonGrid(*, grid_info, size, lat, lon, FT=np.float32, landmask=None, halo=(2, 2, 2)):
    """`LoadLatLonGrid(; grid_info, size, lat, lon, FT=Float32, landmask=nothing,
    halo=(2, 2, 2))` — Grids/lat_lon_grid.jl:230-327: a LatLonGrid from external
    grid information (MITgcm names: RF, RC, DRF, DRC, DXC, DXG, DYC, DYG, RAC,
    hFacW, hFacS, hFacC), halos filled by the x topology (periodic or edge
    copies) and edge copies in y and z; the areas and volumes scaled by hFac."""
    Nx, Ny, Nz = (int(s) for s in size)
    Hx, Hy, Hz = (int(h) for h in halo)
    lat1, lat2 = lat
    lon1, lon2 = lon
    if not (-180 <= lon1 < lon2 <= 180):
        raise AssertionError("-180 <= lon₁ < lon₂ <= 180")
    if not (-80 <= lat1 < lat2 <= 80):
        raise AssertionError("-80 <= lat₁ < lat₂ <= 80")
    periodic = lon1 == -180 and lon2 == 180
    gi = {k: np.asarray(v) for k, v in dict(grid_info).items()}
    Px, Py, P = Nx + 2 * Hx, Ny + 2 * Hy, Nz + 2 * Hz
    f32 = np.float32
    zF = np.zeros(P, f32); zC = np.zeros(P, f32); dzF = np.zeros(P, f32); dzC = np.zeros(P, f32)
    zF[Hz:Nz + Hz + 1] = gi["RF"].astype(f32)                            # :268-271
    zC[Hz:Nz + Hz] = gi["RC"].astype(f32)
    dzF[Hz:Nz + Hz] = gi["DRF"].astype(f32)
    dzC[Hz:Nz + Hz] = gi["DRC"].astype(f32)
    dzC[Hz] *= f32(2.0)                                                  # the first layer has half a cell
    def two(name):
        a = np.zeros((Px, Py), f32, order="F")
        a[Hx:Nx + Hx, Hy:Ny + Hy] = gi[name].astype(f32)
        return a
    dxC, dxF, dyC, dyF, Az = two("DXC"), two("DXG"), two("DYC"), two("DYG"), two("RAC")
    def three(name):
        a = np.ones((Px, Py, P), f32, order="F")
        a[Hx:Nx + Hx, Hy:Ny + Hy, Hz:Nz + Hz] = gi[name].astype(f32)
        return a
    hFW, hFS, hFC = three("hFacW"), three("hFacS"), three("hFacC")
    dzF[0:Hz] = dzF[Hz]; dzC[0:Hz] = dzC[Hz]                              # :281-284
    dzF[Nz + Hz:] = dzF[Nz + Hz - 1]; dzC[Nz + Hz:] = dzC[Nz + Hz - 1]
    zF[0:Hz] = zF[Hz:2 * Hz] + dzF[0] * f32(2)
    zC[0:Hz] = zC[Hz:2 * Hz] + dzC[0] * f32(2)
    zF[Nz + Hz + 1:Nz + 2 * Hz] = zF[Nz + 1:Nz + Hz] - dzF[-1] * f32(2)
    zC[Nz + Hz:Nz + 2 * Hz] = zC[Nz:Nz + Hz] - dzC[-1] * f32(2)
    for a in (dxC, dyC, dxF, dyF, Az):                                   # :291-304
        if periodic:
            a[0:Hx, :] = a[Nx:Nx + Hx, :]
            a[Nx + Hx:, :] = a[Hx:2 * Hx, :]
        else:
            a[0:Hx, :] = a[Hx:Hx + 1, :]
            a[Nx + Hx:, :] = a[Nx + Hx - 1:Nx + Hx, :]
        a[:, 0:Hy] = a[:, Hy:Hy + 1]
        a[:, Ny + Hy:] = a[:, Ny + Hy - 1:Ny + Hy]
    for a in (hFW, hFS, hFC):                                            # :306-316
        if periodic:
            a[0:Hx] = a[Nx:Nx + Hx]
            a[Nx + Hx:] = a[Hx:2 * Hx]
        else:
            a[0:Hx] = a[Hx:Hx + 1]
            a[Nx + Hx:] = a[Nx + Hx - 1:Nx + Hx]
        a[:, 0:Hy] = a[:, Hy:Hy + 1]
        a[:, Ny + Hy:] = a[:, Ny + Hy - 1:Ny + Hy]
        a[:, :, 0:Hz] = a[:, :, Hz:Hz + 1]
        a[:, :, Nz + Hz:] = a[:, :, Nz + Hz - 1:Nz + Hz]
    dz3 = dzF[None, None, :]
    m = dict(dxC=dxC, dyC=dyC, dxF=dxF, dyF=dyF, Az=Az,                   # :318-325
             Ax=np.asfortranarray(((dz3 * dyF[:, :, None]) * hFW).astype(f32)),
             Ay=np.asfortranarray(((dz3 * dxF[:, :, None]) * hFS).astype(f32)),
             Vol=np.asfortranarray(((dz3 * Az[:, :, None]) * hFC).astype(f32)))
    g = LatLonGrid(size=size, lat=lat, lon=lon, z=tuple((float(zF[Hz]), float(zF[Nz + Hz]))), FT=FT,
                   landmask=landmask, halo=halo, _metrics=m)
    g.zF, g.zC, g.dzF, g.dzC = zF, zC, dzF, dzC
    return g
