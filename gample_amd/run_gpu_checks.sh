#!/bin/bash
# GPU check sequence used on the MI355X host: parity tests, smoke, bench,
# kernel-trace profile of a short bench.  Each GPU step has its own time
# limit; a timeout, abort or crash stops the sequence (test failures do not).
set -u
cd "$(dirname "$0")/.."
mkdir -p gpurun_out
step() {  # name, limit, command...
  local name=$1 lim=$2; shift 2
  echo "[step] $name" | tee -a gpurun_out/steps.log
  timeout -k 10 "$lim" "$@" > "gpurun_out/$name.log" 2>&1
  local rc=$?
  echo "[step] $name rc=$rc" | tee -a gpurun_out/steps.log
  if [ $rc -gt 1 ]; then exit $rc; fi
  return 0
}
what=${1:-all}
if [ "$what" = all ] || [ "$what" = tests ]; then
  step gpu_tests 600 python -u -m pytest tests -m gpu -v --timeout 120 --timeout-method thread
fi
if [ "$what" = all ] || [ "$what" = smoke ]; then
  step smoke 180 python -u -c "import __graft_entry__ as g; g.smoke()"
fi
if [ "$what" = all ] || [ "$what" = bench ]; then
  step bench 500 python -u bench.py
fi
if [ "$what" = all ] || [ "$what" = prof ]; then
  export TMPDIR=/tmp
  step prof 400 rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/prof -o bench_kt -- python -u bench.py --no-cpu --no-nn
fi
exit 0
