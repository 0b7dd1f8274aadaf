// Name of libtimestep's rotational velocity-Verlet integrator, as the
// granular_system templates take it (granular_system.h:28-34). The device
// path implements it (kick-drift-kick with accelerations recomputed at the
// drifted state; libgran_amd/csrc/kernels.hip k_kick_drift / k_final_kick).
#ifndef LIBTIMESTEP_ROTATIONAL_INTEGRATOR_H
#define LIBTIMESTEP_ROTATIONAL_INTEGRATOR_H

template <typename field_container_t, typename field_value_t, typename real_t, typename functor_t,
          template <typename, typename> typename step_handler_t>
class rotational_velocity_verlet_half {};

#endif
