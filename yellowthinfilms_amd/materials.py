"""Material and layer types -- the host-side mirror of mod_ModeSolver.jl:21-80.

A material is its 3x3 complex permittivity (or permeability) tensor plus the
kind that decides which branch of _eigen_QV (mod_ModeSolver.jl:203-265) a
layer takes: isotropic / uniaxial eps with isotropic mu use the closed-form
TE/TM modes, anything else the general Berreman eigen-solve.
"""
import numpy as np

EPS_ISOTROPIC, EPS_UNIAXIAL, EPS_BIAXIAL, EPS_GENERAL = 0, 1, 2, 3
MU_ISOTROPIC, MU_GENERAL = 0, 1


class _Material:
    kind = EPS_GENERAL

    def __init__(self, tensor):
        self.eps = np.asarray(tensor, dtype=np.complex128).reshape(3, 3)

    def __getitem__(self, ij):
        return self.eps[ij]

    def __repr__(self):
        return f"{type(self).__name__}({self.eps!r})"


class IsotropicMaterial(_Material):
    """IsotropicMaterial(eps) -- mod_ModeSolver.jl:29-34."""
    kind = EPS_ISOTROPIC

    def __init__(self, eps):
        super().__init__(complex(eps) * np.eye(3))


class UniaxialMaterial(_Material):
    """UniaxialMaterial(eps_t, eps_z), c-axis along z -- mod_ModeSolver.jl:37-46."""
    kind = EPS_UNIAXIAL

    def __init__(self, eps_t, eps_z):
        super().__init__(np.diag([complex(eps_t), complex(eps_t), complex(eps_z)]))


class BiaxialMaterial(_Material):
    """BiaxialMaterial(eps_x, eps_y, eps_z) -- mod_ModeSolver.jl:49-57."""
    kind = EPS_BIAXIAL

    def __init__(self, ex, ey, ez):
        super().__init__(np.diag([complex(ex), complex(ey), complex(ez)]))


class GyrotropicMaterial(_Material):
    """GyrotropicMaterial(eps 3x3) -- mod_ModeSolver.jl:60-65 (any full tensor)."""
    kind = EPS_GENERAL


class FlatOpticalLayer:
    """FlatOpticalLayer(eps_material, d), mu = identity -- mod_ModeSolver.jl:71-74, 108-109."""

    def __init__(self, eps, d):
        self.eps = eps
        self.mu = IsotropicMaterial(1.0)
        self.d = float(d)


class FlatLayer:
    """FlatLayer(eps_material, mu_material, d) -- mod_ModeSolver.jl:76-80."""

    def __init__(self, eps, mu, d):
        self.eps = eps
        self.mu = mu
        self.d = float(d)


def stack_arrays(stack):
    """(eps (n,3,3), mu (n,3,3), d (n,), eps_kind (n,), mu_kind (n,)) of a stack."""
    eps = np.stack([l.eps.eps for l in stack])
    mu = np.stack([l.mu.eps for l in stack])
    d = np.array([l.d for l in stack], dtype=np.float64)
    ek = np.array([l.eps.kind for l in stack], dtype=np.int32)
    mk = np.array([MU_ISOTROPIC if isinstance(l.mu, IsotropicMaterial) else MU_GENERAL for l in stack],
                  dtype=np.int32)
    return eps, mu, d, ek, mk
