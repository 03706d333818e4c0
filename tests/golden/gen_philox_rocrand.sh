#!/bin/bash
# Regenerates tests/golden/philox_rocrand.json from rocRAND's host engine.
set -e
cd "$(dirname "$0")"
/opt/rocm/bin/hipcc -O1 -o /tmp/gen_philox_rocrand gen_philox_rocrand.cpp
/tmp/gen_philox_rocrand > philox_rocrand.json
