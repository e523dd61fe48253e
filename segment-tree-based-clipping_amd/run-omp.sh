#!/bin/sh
# Pairwise clips of the ICPP experiments (the reference's run-omp.sh:3-9) on
# bin/seg of this package.  The dataset (data/poly_data.zip of the reference)
# is not distributed; a line whose files are missing is skipped with a note.
# OMP_NUM_THREADS is kept for the contract; the clip itself runs on the GPU.
cd "$(dirname "$0")"
run() {
    for f in "$@"; do
        if [ ! -r "$f" ]; then
            echo "skipped (missing $f): bin/seg $*"
            return 0
        fi
    done
    OMP_NUM_THREADS=${OMP_NUM_THREADS:-10} bin/seg "$@"
}
run data/s.txt data/c.txt
run data/ne_10m_ocean_0.txt data/continents_521.txt
run data/ne_10m_ocean_0.txt data/continents_1661.txt
run data/lakes_174690.txt data/parks_321571_-226-8_-3-5_lakes_174690_fx.txt
run data/lakes_174690.txt data/parks_34622_-230-6_-5-6_lakes_174690_fx.txt
run data/lakes_174690.txt data/parks_140315_-234-2_-6-8_lakes_174690_fx.txt
run data/lakes_174690.txt data/parks_169840_-230.2_-4.2_lakes_174690_fx.txt
