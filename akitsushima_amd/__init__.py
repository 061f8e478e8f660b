"""akitsushima_amd: an MI355X-native FP64 force engine for akitsushima.

Hand-written HIP kernels for gfx950 behind the reference's force-function
API: Tersoff (one or many species, Tersoff 1989 mixing), Stillinger-Weber
(flat or multi-species records) and alloy EAM (Zhou forms or setfl tables)
energies, forces and virial stress, a device cell-list neighbour list,
batched evaluation of many structures, device-resident NVE, NVT
(Nose-Hoover chain, Langevin, CSVR) and NPT (isotropic or flexible cell)
integrators behind the reference's simulation classes, site energies and the
exact many-body heat flux (Green-Kubo thermal conductivity), a Tersoff
parameter gradient with the reference's fitting interface (parameter set,
loss, residuals), and the batched callers of the force call: force
constants and Gamma-point phonons, and elastic constants.  The compute
path is the HIP library ``libakitsushima_amd.so`` only; importing works
without a GPU, computing does not.
"""
from .neighborlist import NeighbourList
from .potentials import (
    EAMPotential,
    PotentialEngine,
    StillingerWeberPotential,
    TersoffPotential,
    call_force_fn,
    force_fn_from_config,
    make_force_fn,
    tersoff1989_to_triplets,
)
from .setfl import read_setfl, write_setfl
from .site_terms import SiteTerms
from .thermal import ThermalConductivityCalculator
from .simulations import (CSVRSimulation, LangevinSimulation, NPTSimulation, NVESimulation,
                          NVTSimulation)
from .nemd import NEMDSimulation
from .elastic import elastic_constants
from .phonon import force_constants, gamma_frequencies
from .zhou import zhou_eam
from .fitting import (TersoffParameterSet, generic_loss_and_grad, generic_loss_function,
                      generic_residuals, residual_jacobian, residual_length)

__version__ = "0.1.0"

__all__ = [
    "NeighbourList",
    "TersoffPotential",
    "StillingerWeberPotential",
    "EAMPotential",
    "PotentialEngine",
    "make_force_fn",
    "call_force_fn",
    "force_fn_from_config",
    "zhou_eam",
    "read_setfl",
    "SiteTerms",
    "ThermalConductivityCalculator",
    "NEMDSimulation",
    "NVESimulation",
    "NVTSimulation",
    "LangevinSimulation",
    "CSVRSimulation",
    "NPTSimulation",
    "write_setfl",
    "tersoff1989_to_triplets",
    "elastic_constants",
    "force_constants",
    "gamma_frequencies",
    "TersoffParameterSet",
    "generic_loss_function",
    "generic_loss_and_grad",
    "generic_residuals",
    "residual_length",
    "residual_jacobian",
]
