#!/bin/bash
# Builds the reference LB3D (Fortran 90 + MPI) from its own sources, where they
# lie, with amdflang and the image's MPICH (/opt/conda), into oracle/_ref/:
#   oracle/_ref/lb3d_ternary   default build (ternary: red, blue, surfactant)
#   oracle/_ref/lb3d_binary    -DNOSURFACTANT
# Objects and module files go to oracle/_ref/build_<flavour>/. No reference
# file is copied into the repository. Built without XDRF and HDF5 (their
# libraries are absent): checkpoints use checkpoint_format='ori'.
# It is a checker and a CPU baseline only; the product never loads it.
set -euo pipefail
SRC=${LB3D_REF_SRC:-/root/reference/code}
OUT="$(cd "$(dirname "$0")" && pwd)/_ref"
FC=${FC:-amdflang}
MPI=${LB3D_MPI_PREFIX:-/opt/conda}
# module dependency order (code/Makefile)
MODS="lb3d_global lb3d_mpi_parameters lb3d_log lb3d_config lb3d_lattice lb3d_helper lb3d_mpi
lb3d_io_helper lb3d_analysis lb3d_io_hdf5 lb3d_io_xdrf lb3d_force_shanchen lb3d_io_dump_data
lb3d_io_arrstats lb3d_io_stress lb3d_bdist lb3d_bc_leesedwards lb3d_io_checkpoint lb3d_timer lb3d_io
lb3d_random lb3d_invasion lb3d_force_constant lb3d_force_kolmogorov lb3d_force lb3d_bc_periodic
lb3d_bc_bounce lb3d_bc lb3d_ic_rock lb3d_ic lb3d_compatibility lb3d_initialise lb3d_advection
lb3d_collision lb3d_sanity lb3d_time_loop lb3d_finalise lb3d"
[ -d "$SRC" ] || { echo "no reference sources at $SRC" >&2; exit 1; }
command -v "$FC" > /dev/null || { echo "no Fortran compiler $FC" >&2; exit 1; }
[ -f "$MPI/include/mpif.h" ] || { echo "no MPI at $MPI" >&2; exit 1; }
build() {  # flavour, defines
    local dir="$OUT/build_$1"
    mkdir -p "$dir"
    for m in $MODS; do
        "$FC" -cpp -O2 $2 -I"$SRC" -I"$MPI/include" -module-dir "$dir" -c "$SRC/$m.F90" -o "$dir/$m.o" \
            2> "$dir/$m.log" || { cat "$dir/$m.log" >&2; exit 1; }
    done
    local objs=""
    for m in $MODS; do objs="$objs $dir/$m.o"; done
    "$FC" -o "$OUT/lb3d_$1.tmp" $objs -L"$MPI/lib" -lmpifort -lmpi -Wl,-rpath,"$MPI/lib"
    mv "$OUT/lb3d_$1.tmp" "$OUT/lb3d_$1"
}
build ternary ""
build binary "-DNOSURFACTANT"
echo "built $OUT/lb3d_ternary $OUT/lb3d_binary"
