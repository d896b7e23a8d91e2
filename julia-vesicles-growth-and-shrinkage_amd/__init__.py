"""MI355X (gfx950) implementation of the seven-step pseudo-spectral SAV timestep
of Forstylime/Julia-Vesicles-growth-and-shrinkage, exposed through the C ABI of
libsav_hip.so (include/sav_abi.h) and this host-side mirror of the reference's
Config / FieldState / Operators / run_simulation interface."""
from .config import (Config, BDFCoeff, bdf_coeff, set_para_base, stabilisers, SplitMix64, Vesicle, Workload,
                     initial_fields, workload_c1, workload_c2, workload_c3, workload_c4, workload_c5, WORKLOADS,
                     InexactError)
from .sav import (FieldState, Operators, Ensemble, config_c, build_operators, to_spectral, to_physical, update_state,
                  generate_initial_condition, state_from_physical, run_simulation, time_nodes, load_library,
                  save_steps, snapshot_name, write_snapshot, read_snapshot, write_checkpoint, read_checkpoint,
                  resume_simulation, SNAPSHOT_KEYS, initial_fields_3d, initial_fields_2d,
                  SavError, WCNegativeError, EXPORTS, LIB_PATH,
                  SAV_OK, SAV_W_KRYLOV_NOT_CONVERGED, SAV_W_THETA2_NONNEG, SAV_E_ARG, SAV_E_HIP,
                  SAV_E_WC_NEGATIVE, SAV_E_NONFINITE, SAV_E_UNSUPPORTED, SAV_PREC_FP64, SAV_PREC_FP32, SAV_PREC_MIXED, SAV_PLANE_AUTO, SAV_PLANE_REGISTER, SAV_PLANE_LDS, GRAPH_AUTO,
                  SAV_PRECOND_DIAG, SAV_PRECOND_LEARNED, SAV_PRECOND_STENCIL, SAV_PRESENT, SAV_OLD, SAV_NET_NPARAM)
from . import plots  # noqa: E402  (visualize.jl: matplotlib, host side)
