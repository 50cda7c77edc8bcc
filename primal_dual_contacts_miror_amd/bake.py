"""Command-line baking of a reference scene file on the device
(src/Applications/CLIContactSimulation.cpp): loads the YAML scene, runs it to
its baking time and writes <output>.gltf, <output>.npz and <output>.info.

    python -m primal_dual_contacts_miror_amd.bake <input_file> <output_file>
"""
import sys

from . import scene_io

USAGE = """Usage: python -m primal_dual_contacts_miror_amd.bake <input_file> <output_file>
Runs a simulation based on the specified input file and saves results to the output file.

Arguments:
  input_file    Path to the YAML input file describing the simulation scene.
  output_file   Path where the simulation output (GLTF) will be saved.

Options:
  --help        Display this help message and exit.
"""


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) == 1 and argv[0] == "--help":
        print(USAGE)
        return 0
    if len(argv) < 2:
        sys.stderr.write("Error: Missing required arguments.\n")
        print(USAGE)
        return 1
    loaded = scene_io.load_scene(argv[0])
    print("%s: %d bodies, timestep %g, baking time %g" % (argv[0], loaded.scene.n, loaded.params.timestep,
                                                         loaded.baking_time), flush=True)
    times = scene_io.bake(loaded, argv[1])
    print("baked %d frames to t = %g" % (len(times), times[-1] if len(times) else 0.0))
    with open(argv[1].rsplit(".", 1)[0] + ".info") as f:
        print(f.read())
    return 0


if __name__ == "__main__":
    sys.exit(main())
