"""oetqf_amd -- MI355X (gfx950) drop-in for the Oetqf.jl earthquake-cycle hot path.

Host-side mirror of the reference's API (src/Oetqf.jl:29-45): mesh
generation, stress Green's functions, properties, `assemble` and a Tsit5
`solve`, all backed by the HIP library lib/liboetqf_amd.so through its C ABI.
"""
from ._lib import OetqfError, device_count, load as load_library  # noqa: F401
from .ensemble import Ensemble  # noqa: F401
from .group import Group  # noqa: F401
from .equation import (ArrayPartition, Integrator, ODEProblem, ODESolution, Tsit5, VCABM5,  # noqa: F401,E501
                       assemble, assemble_from_mesh, similar, solve)
from .io import SavedState, Val, VΘdεΔ, VΘΔ, wread, wsolve  # noqa: F401
from .greens import (DipSlip, StrikeSlip, get_quadrature, gmsh_quadrature,  # noqa: F401
                     stress_greens_function)
from .mesh import (BEMHex8Mesh, RectOkadaMesh, cosd, gen_gmsh_mesh, gen_mesh,  # noqa: F401
                   gen_mesh_graded, hex8_box, sind)
from .pref import get_matvecmul, set_matvecmul  # noqa: F401
from .property import (CompositePowerLawViscosityProperty, DieterichStateLaw, DilatancyProperty,  # noqa: F401,E501
                       PowerLawViscosityProperty, RateStateQuasiDynamicProperty, dϵ_dt)
