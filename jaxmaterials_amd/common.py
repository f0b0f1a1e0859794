"""Grid description shared by the solver entry points.

Same class name, constructor arguments and attributes as the reference's
``GridSpec`` (src/jaxmaterials/common.py:6-29), so reference scripts that
build one pass it to this package unchanged.
"""

__all__ = ["GridSpec"]


class GridSpec:
    """A periodic box of nx x ny x nz voxels with side lengths Lx, Ly, Lz
    (voxel i along x spans [i Lx/nx, (i+1) Lx/nx))."""

    def __init__(self, nx, ny, nz, Lx, Ly, Lz):
        """Store the voxel counts (nx, ny, nz) and the box lengths (Lx, Ly, Lz)
        along the three axes; fields on this grid are C-ordered arrays of
        shape (nx, ny, nz), z varying fastest."""
        self.nx = nx
        self.ny = ny
        self.nz = nz
        self.Lx = Lx
        self.Ly = Ly
        self.Lz = Lz

    @property
    def number_of_voxels(self):
        """nx * ny * nz."""
        return self.nx * self.ny * self.nz
