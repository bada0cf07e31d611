"""Writes reporting_columns.json: the column names, in order, of the reference's
summary.csv and iterations.csv.

The names are the string literals of writeSummaryHeader and
writeIterationsHeader (src/reporting.cpp:124-166 and :305-328), read from the
source text (nothing of the reference is compiled or run).  Run once in a
container that holds the reference checkout:

    python tests/golden/make_reporting_columns.py /root/reference
"""
import json
import pathlib
import re
import sys


def header_columns(text, function):
    start = text.index(f"void {function}(")
    body = text[start:text.index("\n}", start)]
    literal = "".join(re.findall(r'"((?:[^"\\]|\\.)*)"', body))
    return [name for name in literal.replace("\\n", "").split(",") if name]


def main(reference):
    text = (pathlib.Path(reference) / "src" / "reporting.cpp").read_text()
    data = {
        "about": "Column names of summary.csv and iterations.csv, in order "
                 "(src/reporting.cpp:124-166 and :305-328).",
        "summary": header_columns(text, "writeSummaryHeader"),
        "iterations": header_columns(text, "writeIterationsHeader"),
    }
    out = pathlib.Path(__file__).with_name("reporting_columns.json")
    out.write_text(json.dumps(data, indent=1) + "\n")
    print(f"{out}: {len(data['summary'])} summary, "
          f"{len(data['iterations'])} iteration columns")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
