"""src/read/: the reference's input formats."""
from .collections import Feature, FeatureList, RankedFeatureList

_WS = " \t\n\r\x0b\x0c"


def _lines(path):
    # BufRead::lines(): split on '\n', strip one trailing '\r', no final empty line
    with open(path, "rb") as f:
        data = f.read().decode()
    if not data:
        return []
    parts = data.split("\n")
    if parts[-1] == "":
        parts.pop()
    return [p[:-1] if p.endswith("\r") else p for p in parts]


def read_ranked_feature_list_from_csv(filepath):
    """src/read/read_ranked_feature_list_from_csv.rs:50-70 (raises where it panics)."""
    genes, ranks = [], []
    for i, line in enumerate(_lines(filepath)):
        fields = line.split(",")
        if len(fields) != 2:
            raise ValueError("Invalid format in file %s at line %d" % (filepath, i + 1))
        r = fields[1].strip(_WS)
        digits = r[1:] if r.startswith("+") else r
        if not digits or not digits.isdigit() or not digits.isascii() or int(digits) >= 2**64:
            raise ValueError("Invalid rank value")
        genes.append(Feature(fields[0].strip(_WS)))
        ranks.append(int(digits) & 0xFFFFFFFF)  # `rank as u32`
    return RankedFeatureList.from_(FeatureList(genes), ranks)


def read_feature_list_from_file(filepath):
    """src/read/read_feature_list_from_file.rs:45-53."""
    return FeatureList([Feature(l.strip(_WS)) for l in _lines(filepath)])
