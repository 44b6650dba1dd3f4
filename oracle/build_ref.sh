#!/bin/sh
# Builds oracle/_ref/libmcgb_ref.so: the reference's OWN physics functions, compiled from
# MCnptColloid.c where it lies.  The function bodies are streamed by line range straight into
# gcc (nothing from the reference is written to disk); ref_shim.c (ours) only exports two of the
# reference's static inline helpers.  The whole file cannot be compiled as shipped (syntax
# errors at MCnptColloid.c:367, :3766, :5027, :5185), but these functions compile unchanged:
#   1-365, 372-378   includes, macros, types, globals, prototypes (366-371 is the broken one)
#   1741-1854        GBCollRod
#   1883-2218        find_largest_eigenvalue_index, diagonalize_symmetric_matrix,
#                    compute_liquid_crystal_order_parameters
#   2418-2430        apply_pbc
#   2605-2721        g_omega, pij_GB, compute_dot_product, compute_distance_sq
#   2750-2783        get_system_const
#   2835-3033        initialize_random_positions, initialize_cubic_lattice_positions
#   3050-3078        pto_ransphr
#   3081-3180        print_Trayectoria, print_Qmga, print_pos_xyz
#   3850-4133        error_checker, read_next_parameter, load_initial_conditions
# mpi.h comes from the MPICH installation of this image (only a type in the globals needs it).
# Usage: build_ref.sh <path/to/MCnptColloid.c> <mpi include dir> <output .so>
set -e
SRC="$1"; MPI_INC="$2"; OUT="$3"
HERE=$(cd "$(dirname "$0")" && pwd)
EXPECT=9ff2aea8f6c5dbe720e198c7e09cf3d2c2f80b12a5d1e19b16b1ec40a503b9ed
GOT=$(sha256sum "$SRC" | cut -d' ' -f1)
if [ "$GOT" != "$EXPECT" ]; then
    echo "build_ref.sh: $SRC is not the revision these line ranges were taken from" >&2
    exit 1
fi
mkdir -p "$(dirname "$OUT")"
{
    sed -n '1,365p;372,378p' "$SRC"
    sed -n '1741,1854p;1883,2218p;2418,2430p;2605,2721p;2750,2783p;2835,3033p;3050,3078p;3081,3180p;3850,4133p' "$SRC"
    cat "$HERE/ref_shim.c"
} | gcc -std=gnu11 -O2 -fPIC -shared -w -x c - -I"$MPI_INC" -o "$OUT" -lm
echo "built $OUT"
