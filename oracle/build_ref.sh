#!/usr/bin/env bash
# Build the reference peachgk_md force path (Fortran, serial, -DHF, spline erfc,
# FFTPACK) from its own sources where they lie, plus the C-callable harness in
# oracle/ref_harness.F90, into oracle/_ref/libpgkref.so.
#
# The reference is compiled, never copied: every reference .F90/.F file is read
# from $REF; only generated and built files are written, all under oracle/_ref/.
# The array-size header is regenerated from the reference's config.h with larger
# limits (atoms, FFT grid, cells) so the 100k-atom water box fits; the module
# sources md_global/spme_global/mpi_global are produced from the reference's
# .inc files by the C preprocessor exactly as the reference's makefile does
# (cpp | strip blank and '#' lines), but against the regenerated header.
#
# Usage: oracle/build_ref.sh [REFDIR]   (default /root/reference/peachgk-md)
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${1:-/root/reference/peachgk-md}"
OUT="$HERE/_ref"
FC="${FC:-/opt/rocm/lib/llvm/bin/flang}"
JOBS="${JOBS:-8}"

if [ ! -f "$REF/calforce_hf.F90" ]; then
  echo "build_ref: reference sources not found at $REF" >&2
  exit 2
fi
if [ ! -x "$FC" ]; then
  echo "build_ref: flang not found at $FC" >&2
  exit 2
fi

rm -rf "$OUT/obj" "$OUT/mod" "$OUT/gen" "$OUT/cfg"
mkdir -p "$OUT/obj" "$OUT/mod" "$OUT/gen" "$OUT/cfg"

# --- regenerated size header -------------------------------------------------
sed -e 's/^#define _MAXNATOM_ .*/#define _MAXNATOM_ 110000/' \
    -e 's/^#define _MAXCELLX_ .*/#define _MAXCELLX_ 16/' \
    -e 's/^#define _MAXCELLY_ .*/#define _MAXCELLY_ 16/' \
    -e 's/^#define _MAXCELLZ_ .*/#define _MAXCELLZ_ 32/' \
    -e 's/^#define _MAXNHFREGION_ .*/#define _MAXNHFREGION_ 8/' \
    -e 's/^#define _MAXFFT1_ .*/#define _MAXFFT1_ 128/' \
    -e 's/^#define _MAXFFT2_ .*/#define _MAXFFT2_ 128/' \
    -e 's/^#define _MAXFFT3_ .*/#define _MAXFFT3_ 192/' \
    -e 's/^#define _MAX_GRID_ .*/#define _MAX_GRID_ 192/' \
    "$REF/config.h" > "$OUT/cfg/config.h"

# --- module sources from the .inc files (stdin so "config.h" resolves to cfg/)
for m in md_global spme_global mpi_global; do
  (cd "$OUT/cfg" && cpp -P - < "$REF/$m.inc" | awk '!/^ *$/ && !/^#/' > "$OUT/gen/$m.F90")
done

FFLAGS="-O2 -fPIC -cpp -DHF -D_CSTMNB_V1 -module-dir $OUT/mod -I$OUT/mod"

compile() {  # src obj
  "$FC" $FFLAGS -c "$1" -o "$2" 2> "$2.log" || { cat "$2.log" >&2; return 1; }
}

# modules first, in dependency order
compile "$OUT/gen/md_global.F90"   "$OUT/obj/md_global.o"
compile "$OUT/gen/mpi_global.F90"  "$OUT/obj/mpi_global.o"
compile "$OUT/gen/spme_global.F90" "$OUT/obj/spme_global.o"
for m in time_global interface_tools interface_timer interface_interact \
         interface_preproc interface_mdtech interface_mddriver cstmnb SFMT; do
  compile "$REF/$m.F90" "$OUT/obj/$m.o"
done

# every other translation unit of the scalar build except the program itself
SKIP=" md_global spme_global mpi_global time_global interface_tools interface_timer
       interface_interact interface_preproc interface_mdtech interface_mddriver cstmnb
       SFMT main_md "
MPIONLY=" fft_pme_mpi ewaldk_mpi ewaldk_mpi_hf rattle_mpi rattle_mtk_mpi shake_mpi
         MPIsum_all_atm_heatf MPIsum_all_pot_for MPIsum_all_viri "
SRCS=()
for f in "$REF"/*.F90 "$REF"/*.F; do
  b="$(basename "$f")"; b="${b%.*}"
  case " $(echo $SKIP $MPIONLY) " in *" $b "*) continue;; esac
  SRCS+=("$f")
done
export -f compile; export FC FFLAGS OUT
printf '%s\n' "${SRCS[@]}" | xargs -P "$JOBS" -I{} bash -c \
  'f="{}"; b="$(basename "$f")"; compile "$f" "$OUT/obj/${b%.*}.o"'

compile "$HERE/ref_harness.F90" "$OUT/obj/zz_ref_harness.o"

"$FC" -shared -fPIC -o "$OUT/libpgkref.so" "$OUT"/obj/*.o

# The same reference with its own per-term timers on (-DTIME_M in calforce and
# calforce_hf, as the reference's makefiles allow): the CPU baseline's split.
mkdir -p "$OUT/timed"
for f in calforce calforce_hf; do
  "$FC" $FFLAGS -DTIME_M -c "$REF/$f.F90" -o "$OUT/timed/$f.o" 2> "$OUT/timed/$f.log" ||
    { cat "$OUT/timed/$f.log" >&2; exit 1; }
done
TOBJS=()
for o in "$OUT"/obj/*.o; do
  case "$(basename "$o")" in
    calforce.o|calforce_hf.o) ;;
    *) TOBJS+=("$o");;
  esac
done
"$FC" -shared -fPIC -o "$OUT/libpgkref_timed.so" "${TOBJS[@]}" "$OUT/timed/calforce.o" \
  "$OUT/timed/calforce_hf.o"
rm -f "$OUT"/timed/*.log

# The same reference with the drop-in shim (thrid_amd/fortran/calforce_gpu.F90)
# in place of calforce.o, calforce_hf.o, mklist.o and mklist_cell.o, linked to
# the MI355X library: the link itself checks that the shim satisfies every
# caller of those four routines.
PKG="$HERE/../thrid_amd"
if [ -f "$PKG/libpgkforce.so" ]; then
  mkdir -p "$OUT/shim"
  "$FC" -O2 -fPIC -cpp -DHF -module-dir "$OUT/shim" -I"$OUT/mod" \
    -c "$PKG/fortran/pgk_force_mod.F90" -o "$OUT/shim/pgk_force_mod.o"
  "$FC" -O2 -fPIC -cpp -DHF -module-dir "$OUT/shim" -I"$OUT/mod" -I"$OUT/shim" \
    -c "$PKG/fortran/calforce_gpu.F90" -o "$OUT/shim/calforce_gpu.o"
  OBJS=()
  for o in "$OUT"/obj/*.o; do
    case "$(basename "$o")" in
      calforce.o|calforce_hf.o|mklist.o|mklist_cell.o) ;;
      *) OBJS+=("$o");;
    esac
  done
  "$FC" -shared -fPIC -o "$OUT/libpgkref_gpu.so" "${OBJS[@]}" "$OUT/shim/pgk_force_mod.o" \
    "$OUT/shim/calforce_gpu.o" -L"$PKG" -lpgkforce -Wl,-rpath,'$ORIGIN/../../thrid_amd' \
    -Wl,--no-undefined -Wl,--allow-shlib-undefined
  echo "build_ref: built $OUT/libpgkref_gpu.so"
fi
rm -f "$OUT"/obj/*.log
echo "build_ref: built $OUT/libpgkref.so"
