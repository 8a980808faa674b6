#!/usr/bin/env bash
# Profiles the bench workload with rocprofv3: kernel trace + stats, then HBM byte counters
# (FETCH_SIZE, WRITE_SIZE) in separate passes, each also for the calibration copy kernel, then
# writes summary.md / summary.json (prof_summary.py) next to the raw results.
# Usage (from the repository root, on a GPU box): bash staticbtesolver_amd/profile.sh [bench args]
# (results under $PROF_OUT, default gpurun_out/prof; PROF_KERNEL=<name part> reports the most
# launched instantiation of that kernel instead of the busiest kernel)
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
OUT=${PROF_OUT:-$ROOT/gpurun_out/prof}
mkdir -p "$OUT"
export TMPDIR=/tmp
ARGS=${*:-"--profile-run"}
cd /tmp
echo "== kernel trace: bench.py $ARGS"
timeout -k 10 300 rocprofv3 --kernel-trace --stats -d "$OUT/trace" -o bench -- python3 "$ROOT/bench.py" $ARGS
# with PROF_KERNEL, counters are collected for that kernel only (collecting them serialises every
# dispatch, which makes a run with thousands of small launches very slow)
FILTER=()
[ -n "${PROF_KERNEL:-}" ] && FILTER=(--kernel-include-regex "$PROF_KERNEL")
for C in FETCH_SIZE WRITE_SIZE; do
  echo "== pmc $C: bench.py $ARGS"
  timeout -s KILL 200 rocprofv3 --pmc $C "${FILTER[@]}" -d "$OUT/pmc_$C" -o bench -- python3 "$ROOT/bench.py" $ARGS
  echo "== pmc $C: hbm_copy 1024 MiB"
  timeout -s KILL 60 rocprofv3 --pmc $C -d "$OUT/pmc_copy_$C" -o copy -- "$ROOT/staticbtesolver_amd/bin/hbm_copy" 1024 3
done
echo "== summary"
python3 "$ROOT/staticbtesolver_amd/prof_summary.py" "$OUT" "${TITLE:-bench.py $ARGS}" \
  "${PROF_KERNEL:+PROF_KERNEL=$PROF_KERNEL }bash staticbtesolver_amd/profile.sh $ARGS" 1024 "$OUT/summary.json" > "$OUT/summary.md"
echo "== done"
