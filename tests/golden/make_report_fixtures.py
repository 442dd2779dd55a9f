"""Write the report and input fixtures from the reference's files (read as
text, nothing executed):

  report_formats.json   the printf formats of printStats, printRange and
                        storeHistogram and the printPopulation labels
                        (src/abms.c), with the lines they come from;
  temp_sample.txt       the first 48 lines of temp.txt (input data);
  environment_sample.txt the first 48 lines of environment.txt (input data).

Run in the container that has the reference mounted:
    python tests/golden/make_report_fixtures.py /root/reference
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def line_of(text, pos):
    return text[:pos].count("\n") + 1


def main(ref):
    abms = os.path.join(ref, "src", "abms.c")
    text = open(abms).read()
    out = {}
    m = re.search(r"void printStats\(.*?\{(.*?)\n\}", text, re.S)
    body = m.group(1)
    fmts = re.findall(r'^\s*printf\("((?:[^"\\]|\\.)*)"', body, re.M)
    out["printStats"] = {"formats": [f.encode().decode("unicode_escape") for f in fmts],
                         "source": f"src/abms.c:{line_of(text, m.start())}"}
    m = re.search(r'void printRange\(.*?printf\("((?:[^"\\]|\\.)*)"', text, re.S)
    out["printRange"] = {"format": m.group(1).encode().decode("unicode_escape"),
                         "source": f"src/abms.c:{line_of(text, m.start())}"}
    m = re.search(r"void printPopulation\(.*?\{(.*?)\n\}", text, re.S)
    labels = re.findall(r'printf\("((?:[^"\\]|\\.)*)"\); printRange', m.group(1))
    out["printPopulation"] = {"labels": [x.encode().decode("unicode_escape") for x in labels],
                              "source": f"src/abms.c:{line_of(text, m.start())}"}
    m = re.search(r'fprintf\(hFile, "((?:[^"\\]|\\.)*)"', text)
    out["histogram"] = {"format": m.group(1).encode().decode("unicode_escape"),
                        "source": f"src/abms.c:{line_of(text, m.start())}"}
    for name in ("temp.txt", "environment.txt"):
        lines = open(os.path.join(ref, name)).read().splitlines()
        out[name] = {"lines": len(lines)}
        with open(os.path.join(HERE, name.replace(".txt", "_sample.txt")), "w") as f:
            f.write("\n".join(lines[:48]) + "\n")
    with open(os.path.join(HERE, "report_formats.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
