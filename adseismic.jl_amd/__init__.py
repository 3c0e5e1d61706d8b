"""MI355X-native wave-propagation hot path of ADSeismic.jl (acoustic PML and elastic
staggered-grid time steppers, hand-written adjoints, batched checkpointed FWI gradients).

All compute runs in ``lib/libadseismic_hip.so`` (HIP, gfx950) through the C-ABI declared in
``include/adseismic.h``; this package is the host-side mirror of the reference's Julia API.
The directory name is not a Python identifier: load it with ``__graft_entry__.load_package()``
(registered as the module ``adseismic_amd``).
"""
from ._lib import ADSError, load as load_library  # noqa: F401
from .params import (AcousticPropagatorParams, AcousticReceiver, AcousticSource,  # noqa: F401
                     ElasticPropagatorParams, ElasticReceiver, ElasticSource, Geometry,
                     MPIElasticPropagatorParams, MPIElasticReceiver, MPIElasticSource,
                     compute_PML_Params_mpi)
from .acoustic import (AcousticPlan, AcousticPropagator, AcousticPropagatorSolver,  # noqa: F401
                       SimulatedObservation_, acoustic_one_step, acoustic_one_step_customop,
                       acoustic_traces, check_cfl, compute_PML_Params_, getid2)
from .elastic import (ElasticPlan, ElasticPropagator, ElasticPropagatorSolver,  # noqa: F401
                      MPIElasticPropagator, MPIElasticPropagatorSolver, MPISimulatedObservation_,
                      SimulatedObservation_elastic, elastic_traces)
from .elastic import compute_PML_Params as compute_PML_Params_elastic  # noqa: F401
from .utils import (Gauss, Ricker, compute_lame_parameters, compute_properties,  # noqa: F401
                    ricker_freq)
from .io_mat import (load_acoustic_model, load_acoustic_receiver, load_acoustic_source,  # noqa: F401
                     load_elastic_model, load_elastic_receiver, load_elastic_source, load_params,
                     matread, safe_vec)
from .fwi import (AcousticFWI, Adam_, LBFGS_, acoustic_gradient_in_batches,  # noqa: F401
                  acoustic_traces_in_batches, compute_loss_and_grads)
from .nn import Generator, UpsamplingGenerator, add_initial_model  # noqa: F401
from .mpi_acoustic import (MPIAcousticPropagator, MPIAcousticPropagatorParams,  # noqa: F401
                           MPIAcousticPropagatorSolver, MPIAcousticReceiver, MPIAcousticSource,
                           MPISimulatedObservation_acoustic, compute_PML_Params_mpi_acoustic,
                           mpi_acoustic_one_step)

SimulatedObservation = SimulatedObservation_
