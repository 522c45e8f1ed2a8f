#!/usr/bin/env bash
# Compiles the reference 2DENSE Fortran sources where they lie under
# /root/reference/source with AMD flang (the reference's own CMake hard-codes
# gfortran, which this image lacks) into oracle/_ref/:
#   2dense_ref          the unmodified reference, -O2 -fopenmp
#   2dense_ref_racefix  the same sources with private(char_var) added to the
#                       four eta-sweep OpenMP regions that race on it
#                       (scheme_weno5th_CH.f90:215,301; scheme_weno5th_SC.f90:435,542),
#                       used only for multi-thread timing of switches 104/105.
# The patched copy lives in a temporary directory outside the repository and
# is deleted after the build; only the binaries land in oracle/_ref/.
set -euo pipefail
SRC=${REF_SRC:-/root/reference/source}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
if [ ! -d "$SRC" ]; then
  echo "reference sources not found at $SRC; skipping reference build" >&2
  exit 0
fi
mkdir -p "$OUT"
order() {  # module files first, then everything else
  local d=$1
  echo "$d/utility/var_index.f90 $d/utility/flow_mod_general.f90 $d/inv_scheme/weno_mod.f90"
  ls "$d"/*/*.f90 | grep -v -e '/var_index.f90$' -e '/flow_mod_general.f90$' -e '/weno_mod.f90$'
}
TMP=$(mktemp -d /tmp/dense_ref_build.XXXXXX)
trap 'rm -rf "$TMP"' EXIT
( cd "$TMP" && mkdir plain && cd plain && \
  $FC -O2 -fopenmp -w -o "$OUT/2dense_ref" $(order "$SRC") )
# race-fixed variant
mkdir -p "$TMP/src" && cp -r "$SRC"/. "$TMP/src/"
python3 - "$TMP/src" <<'PY'
import sys, re, pathlib
root = pathlib.Path(sys.argv[1]) / "inv_scheme"
# Add private(char_var) to every OpenMP region in the eta sweeps that lacks it.
for name in ("scheme_weno5th_CH.f90", "scheme_weno5th_SC.f90"):
    p = root / name
    lines = p.read_text().split("\n")
    out = []
    i = 0
    while i < len(lines):
        ln = lines[i]
        out.append(ln)
        if ln.strip().upper().startswith("!$OMP PARALLEL DO"):
            # collect the directive block
            j = i + 1
            block = [ln]
            while j < len(lines) and lines[j].strip().lower().startswith("!$omp") and not lines[j].strip().upper().startswith("!$OMP END"):
                block.append(lines[j]); out.append(lines[j]); j += 1
            text = " ".join(block).lower()
            k = j
            while k < len(lines) and not lines[k].strip().upper().startswith("!$OMP END PARALLEL"):
                k += 1
            if "char_var" not in text and any("char_var" in l.lower() for l in lines[j:k]):
                # last directive line must continue
                last = out.pop()
                out.append(last.rstrip() + " &")
                out.append("!$omp private(char_var)")
            i = j
            continue
        i += 1
    p.write_text("\n".join(out))
PY
( cd "$TMP" && mkdir fixed && cd fixed && \
  $FC -O2 -fopenmp -w -o "$OUT/2dense_ref_racefix" $(order "$TMP/src") )
echo "built: $OUT/2dense_ref $OUT/2dense_ref_racefix"
