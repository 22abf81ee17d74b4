!< Prints the storage size of each bind(C) type of sp_hip_binding; the sizes
!< must equal the C struct sizes of include/sp_hip.h.
program abi_check
    use iso_c_binding
    use sp_hip_binding
    implicit none
    type(sp_config) :: c
    type(sp_mover_stats) :: m
    type(sp_local_dist) :: l
    type(sp_dist_config) :: d
    type(sp_particles) :: p
    type(sp_variant) :: v
    type(sp_tune_result) :: t
    print "(7I6)", c_sizeof(c), c_sizeof(m), c_sizeof(l), c_sizeof(d), c_sizeof(p), c_sizeof(v), &
        c_sizeof(t)
end program abi_check
