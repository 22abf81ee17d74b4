!*******************************************************************************
!< Prints the MHD time stamps the driver would use, one per line:
!<   tstamps_check <dir_mhd_data> <t_start> <t_end> <tmax_mhd>
!< reads dir_mhd_data/time_stamps.dat with load_tstamps_mhd (the -vdt path of
!< the driver).  A check of the reader that needs no GPU.
!*******************************************************************************
program tstamps_check
    use gpat_hip, only: init_tstamps_mhd, load_tstamps_mhd, tstamps_mhd
    implicit none
    character(len=256) :: dir, arg
    integer :: t_start, t_end, tmax_mhd, i
    call get_command_argument(1, dir)
    call get_command_argument(2, arg)
    read(arg, *) t_start
    call get_command_argument(3, arg)
    read(arg, *) t_end
    call get_command_argument(4, arg)
    read(arg, *) tmax_mhd
    call init_tstamps_mhd(t_start, t_end)
    call load_tstamps_mhd(t_start, t_end, tmax_mhd, trim(dir))
    do i = 1, size(tstamps_mhd)
        write(*, "(ES24.16)") tstamps_mhd(i)
    enddo
end program tstamps_check
