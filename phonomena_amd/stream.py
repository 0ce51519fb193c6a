"""Practical HBM bandwidth of a read/write mix (pf_stream_probe).

    python -m phonomena_amd.stream READS WRITES N [REPS]

Streams READS input and WRITES output float arrays of N elements through the
device with a plain coalesced kernel of known traffic and prints one JSON
line: GB/s, ms per pass and the bytes one pass moves.  Used for the
bandwidth figure beside the vendor peak and for calibrating FETCH_SIZE /
WRITE_SIZE counters (a pass moves exactly READS*N*4 bytes in, WRITES*N*4 out).
"""
import json
import sys

from phonomena_amd import fdtd


def main(argv):
    if len(argv) < 3:
        print(__doc__, file=sys.stderr)
        return 2
    r, w, n = int(argv[0]), int(argv[1]), int(argv[2])
    reps = int(argv[3]) if len(argv) > 3 else 20
    gbs, ms = fdtd.stream_probe(r, w, n, reps)
    print(json.dumps({"reads": r, "writes": w, "n": n, "gbs": gbs, "ms": ms,
                      "bytes_in": r * n * 4, "bytes_out": w * n * 4}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
