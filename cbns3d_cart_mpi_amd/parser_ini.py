"""`.ini` -> `input.txt` converter (the reference's python/parser_ini.py role).

    python -m cbns3d_cart_mpi_amd.parser_ini taylor_green.ini [input.txt]

Writes every value of every section, in file order, one per line: the
positional layout that read_input consumes (src/read_input.cpp:24-72).  Keys
keep their case.  Prints "key = value" for each entry, as the reference's
script does.
"""
import configparser
import sys


def convert(ini_path, out_path="input.txt", echo=True):
    cfg = configparser.ConfigParser()
    cfg.optionxform = str
    if not cfg.read(ini_path):
        raise FileNotFoundError(ini_path)
    values = []
    for sec in cfg.sections():
        for key in cfg[sec]:
            if echo:
                print(f"{key} = {cfg[sec][key]}")
            values.append(cfg[sec][key])
    with open(out_path, "w") as f:
        for v in values:
            f.write(v + "\n")
    return values


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m cbns3d_cart_mpi_amd.parser_ini <file.ini> [input.txt]", file=sys.stderr)
        return 2
    convert(argv[0], argv[1] if len(argv) > 1 else "input.txt")
    return 0


if __name__ == "__main__":
    sys.exit(main())
