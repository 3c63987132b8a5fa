#!/bin/bash
# rocprof_passes.sh -- profile the DFSPH benchmark with rocprofv3.
#
#   bash splishsplash_for_cuda_amd/rocprof_passes.sh OUTDIR [bench.py arguments...]
#
# One kernel-trace + stats run, then one run per hardware counter (each in a
# run of its own, as rocprofv3 does not split counters over passes), and the
# same counter runs over the plain copy kernel of known size used to calibrate
# FETCH_SIZE / WRITE_SIZE.  EXTRA_COUNTERS adds passes: a space-separated list
# whose entries join the counters of one pass with '+' (within the per-block
# limits of one pass).  TRACE_ARGS (default: the same arguments) gives the
# kernel-trace run its own bench.py arguments, e.g. the default bench command
# while the counter passes run a shorter one.  Every step has its own time
# limit and the script stops at the first failure.  Summaries:
# python splishsplash_for_cuda_amd/rocprof_summary.py OUTDIR
set -u
OUT=${1:?output directory}
shift
BENCH_ARGS=("$@")
[ ${#BENCH_ARGS[@]} -eq 0 ] && BENCH_ARGS=(--steps 5 --warmup 5 --no-cpu-baseline)
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
mkdir -p "$OUT"
OUT=$(cd "$OUT" && pwd)
export TMPDIR=/tmp
cd "$ROOT"

if [ -n "${TRACE_ARGS+x}" ]; then read -r -a T_ARGS <<< "$TRACE_ARGS"; else T_ARGS=("${BENCH_ARGS[@]}"); fi
echo "== kernel trace: bench.py ${T_ARGS[*]}"
timeout -k 10 400 rocprofv3 --kernel-trace --stats --output-format csv -d "$OUT/trace" -o run -- \
    python3 bench.py "${T_ARGS[@]}" > "$OUT/trace.log" 2>&1 || { echo "trace failed"; tail -5 "$OUT/trace.log"; exit 1; }

for m in FETCH_SIZE WRITE_SIZE; do
    echo "== copy calibration $m"
    timeout -s KILL 90 rocprofv3 --pmc $m --output-format csv -d "$OUT/copy_$m" -o copy -- \
        "$HERE/copy_calibration" > "$OUT/copy_$m.log" 2>&1 || { echo "copy $m failed"; exit 1; }
done

for m in FETCH_SIZE WRITE_SIZE ${EXTRA_COUNTERS:-}; do
    echo "== bench $m"
    timeout -s KILL 240 rocprofv3 --pmc ${m//+/ } --output-format csv -d "$OUT/bench_$m" -o bench -- \
        python3 bench.py "${BENCH_ARGS[@]}" > "$OUT/bench_$m.log" 2>&1 || { echo "bench $m failed"; tail -5 "$OUT/bench_$m.log"; exit 1; }
done
echo "done: $OUT"
