"""MI355X-native VMC inner loop for the spinful Fermi-Hubbard model.

Reference-shaped host classes (sampler, models, Hamiltonian, energy /
gradient functions, SR preconditioners, VMC driver) over the HIP kernels of
libvmcx.so.  The library must be built (python -m vmc_torch_amd.build);
nothing here falls back to the CPU.
"""
from .hamiltonian import SquareLattice, SpinfulFermion, spinful_Fermi_Hubbard_square_lattice_torch
from .models import (WavefunctionModel_GPU, SlaterDeterminant_GPU, NNBF_GPU,
                     fPEPS_Model_GPU, fPEPS_Model_reuse_GPU, NNfTNS_Model_GPU)
from .sampler import (SamplerGPU, MetropolisExchangeSpinfulSamplerGPU,
                      MetropolisExchangeSpinfulSamplerReuse_GPU)
from .vmc_utils import (random_initial_config, initialize_walkers, evaluate_energy,
                        evaluate_energy_reuse, compute_grads_gpu)

__all__ = [
    "SquareLattice", "SpinfulFermion", "spinful_Fermi_Hubbard_square_lattice_torch",
    "WavefunctionModel_GPU", "SlaterDeterminant_GPU", "NNBF_GPU", "fPEPS_Model_GPU",
    "fPEPS_Model_reuse_GPU", "NNfTNS_Model_GPU", "SamplerGPU",
    "MetropolisExchangeSpinfulSamplerGPU", "MetropolisExchangeSpinfulSamplerReuse_GPU",
    "random_initial_config", "initialize_walkers", "evaluate_energy", "evaluate_energy_reuse",
    "compute_grads_gpu",
]
