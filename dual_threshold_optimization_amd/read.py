"""src/read/: the reference's input formats."""
from .collections import Feature, FeatureList, RankedFeatureList

# Unicode White_Space (PropList.txt): what Rust's str::trim strips.  Python's
# str.strip() with no argument also strips U+001C..U+001F, which Rust keeps.
_WS = ("\t\n\x0b\x0c\r \x85\xa0\u1680\u2000\u2001\u2002\u2003\u2004\u2005\u2006\u2007\u2008"
       "\u2009\u200a\u2028\u2029\u202f\u205f\u3000")
_BAD_UTF8 = 'Could not read line: Error { kind: InvalidData, message: "stream did not contain valid UTF-8" }'


def _lines(path):
    """BufRead::lines(): split on '\\n', strip one trailing '\\r', no final
    empty line.  Yields each line's bytes; decoding (which fails on invalid
    UTF-8, as lines() does) is the caller's, line by line, in order."""
    with open(path, "rb") as f:
        data = f.read()
    if not data:
        return []
    parts = data.split(b"\n")
    if parts[-1] == b"":
        parts.pop()
    return [p[:-1] if p.endswith(b"\r") else p for p in parts]


def _decode(raw):
    try:
        return raw.decode("utf-8")
    except UnicodeDecodeError:
        raise ValueError(_BAD_UTF8) from None


def read_ranked_feature_list_from_csv(filepath):
    """src/read/read_ranked_feature_list_from_csv.rs:50-70 (raises where it panics)."""
    genes, ranks = [], []
    for i, raw in enumerate(_lines(filepath)):
        fields = _decode(raw).split(",")
        if len(fields) != 2:
            raise ValueError("Invalid format in file %s at line %d" % (filepath, i + 1))
        r = fields[1].strip(_WS)
        digits = r[1:] if r.startswith("+") else r
        if not digits or not digits.isascii() or not digits.isdigit() or int(digits) >= 2**64:
            raise ValueError("Invalid rank value")
        genes.append(Feature(fields[0].strip(_WS)))
        ranks.append(int(digits) & 0xFFFFFFFF)  # `rank as u32`
    return RankedFeatureList.from_(FeatureList(genes), ranks)


def read_feature_list_from_file(filepath):
    """src/read/read_feature_list_from_file.rs:45-53."""
    return FeatureList([Feature(_decode(raw).strip(_WS)) for raw in _lines(filepath)])
