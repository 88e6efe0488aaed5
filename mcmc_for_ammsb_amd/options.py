"""Options of the sampler, mirroring ``mcmc::Options``.

Reference: c++/src/mcmc/options.h:129-315 (Boost program_options with dotted
``--mcmc.*`` flags, short aliases and an optional ``--config`` file) and the
strategy names of c++/src/mcmc/types.h:11-60.  Defaults and the derived
defaults (``a = b^-c``, ``alpha = 1/K``, options.h:257-262) are the same;
unrecognised arguments are kept in ``remains`` like ``allow_unregistered``
(options.h:244-255).  Integers parse as the reference's ``parse<T>`` (0x hex,
leading-0 octal, k/m/g suffixes; c++/src/mcmc/options.h:39-57, options.cc:5-67).
"""
import ctypes
import dataclasses
import math

from . import _lib

STRATEGIES = {
    "stratified-random-node": _lib.STRATIFIED_RANDOM_NODE,
    "stratified-random-node-links": _lib.STRATIFIED_RANDOM_NODE_LINKS,
    "stratified-random-node-nonlinks": _lib.STRATIFIED_RANDOM_NODE_NONLINKS,
    "random-edge": _lib.RANDOM_EDGE,
}


def _parse_size_t(arg):
    """parse_size_t (options.cc:5-67): "0x" hexadecimal, a leading "0" octal
    (its digit test accepts 8 and 9, as there), else decimal; the rest of
    the string may be one suffix k/kb, m/mb, g/gb (any case) for 2^10,
    2^20, 2^30."""
    n, base, i = 0, 10, 0
    if arg[:2] in ("0x", "0X"):
        base, i = 16, 2
    elif arg[:1] == "0":
        base, i = 8, 1
    while i < len(arg):
        ch = arg[i]
        if base == 16 and ch in "0123456789abcdefABCDEF":
            n = n * 16 + int(ch, 16)
        elif base <= 10 and ch.isdigit() and ch.isascii():
            n = n * base + (ord(ch) - ord("0"))
        else:
            rest = arg[i:].lower()
            shift = {"g": 30, "gb": 30, "m": 20, "mb": 20, "k": 10, "kb": 10}.get(rest)
            if shift is None:
                raise ValueError(f"Unknown characters in number: {ch!r}")
            n <<= shift
            break
        if n >= 1 << 64:
            raise ValueError("Overflow in parse_size_t")
        i += 1
    if n >= 1 << 64:
        raise ValueError("Overflow in parse_size_t")
    return n


def _parse_integral(text, lo, hi):
    """mcmc::parse<T> (options.h:39-57): a leading '-' negates."""
    s = str(text)
    if s.startswith("-"):
        v = -_parse_size_t(s[1:])
        if v < lo:
            raise ValueError("number too small for type")
        return v
    v = _parse_size_t(s)
    if v > hi:
        raise ValueError("number too large for type")
    return v


def parse_real(text):
    """A floating-point option, read as lexical_cast<double> reads it: the
    whole text one decimal number (sign, digits, point, exponent) or
    inf/nan; no blanks, no hexadecimal form, no digit separators."""
    s = str(text)
    if not s or any(ch.isspace() or ch in "_xX(" for ch in s):
        raise ValueError(f"invalid number {s!r}")
    return float(s)


def parse_size(text):
    """A size_t option (no negative values)."""
    return _parse_integral(text, 0, (1 << 63) - 1)


def parse_int(text):
    """An int option."""
    return _parse_integral(text, -(1 << 31), (1 << 31) - 1)


@dataclasses.dataclass
class Options:
    # --mcmc.* (options.h:139-189)
    alpha: float = 0.0
    eta0: float = 1.0
    eta1: float = 1.0
    epsilon: float = 0.0000001
    a: float = 0.0
    b: float = 1024.0
    c: float = 0.5
    K: int = 300
    mini_batch_size: int = 0
    num_node_sample: int = 0
    strategy: str = "stratified-random-node"
    sampler_max_source: int = 1
    max_iteration: int = 10000000
    interval: int = 0
    stats_print_interval: int = 0
    num_updates: int = 1000
    held_out_ratio: float = 0.0
    random_seed: int = 42
    convergence_threshold: float = 0.000000000001
    # input options (options.h:192-205)
    input_filename: str = ""
    input_class: str = "relativity"
    input_contiguous: bool = False
    input_compressed: bool = False
    # this implementation
    precision: str = "fp64"
    device: int = 0
    verbose: bool = False
    remains: list = dataclasses.field(default_factory=list)

    # flag name -> (field, converter); short aliases of options.h
    _FLAGS = {
        "mcmc.alpha": ("alpha", parse_real), "mcmc.eta0": ("eta0", parse_real),
        "mcmc.eta1": ("eta1", parse_real), "mcmc.epsilon": ("epsilon", parse_real),
        "mcmc.a": ("a", parse_real), "mcmc.b": ("b", parse_real), "mcmc.c": ("c", parse_real),
        "mcmc.K": ("K", parse_size), "K": ("K", parse_size),
        "mcmc.mini-batch-size": ("mini_batch_size", parse_size), "m": ("mini_batch_size", parse_size),
        "mcmc.num-node-sample": ("num_node_sample", parse_size), "n": ("num_node_sample", parse_size),
        "mcmc.strategy": ("strategy", str),
        "mcmc.sampler.max-source": ("sampler_max_source", parse_size),
        "mcmc.max-iteration": ("max_iteration", parse_size), "x": ("max_iteration", parse_size),
        "mcmc.interval": ("interval", parse_size), "i": ("interval", parse_size),
        "mcmc.stats": ("stats_print_interval", parse_size), "I": ("stats_print_interval", parse_size),
        "mcmc.num-updates": ("num_updates", parse_size),
        "mcmc.held-out-ratio": ("held_out_ratio", parse_real),
        "h": ("held_out_ratio", parse_real),
        "mcmc.seed": ("random_seed", parse_int),
        "mcmc.convergence": ("convergence_threshold", parse_real),
        "mcmc.input.file": ("input_filename", str), "f": ("input_filename", str),
        "mcmc.input.class": ("input_class", str), "c": ("input_class", str),
        "mcmc.precision": ("precision", str),
        "mcmc.device": ("device", parse_int),
    }
    _ALIASES = {"K": "mcmc.K", "m": "mcmc.mini-batch-size", "n": "mcmc.num-node-sample",
                "x": "mcmc.max-iteration", "i": "mcmc.interval", "I": "mcmc.stats",
                "h": "mcmc.held-out-ratio", "f": "mcmc.input.file", "c": "mcmc.input.class"}
    _SWITCHES = {"mcmc.input.contiguous": "input_contiguous",
                 "mcmc.input.compressed": "input_compressed", "verbose": "verbose"}

    @classmethod
    def parse(cls, argv):
        """Parse ``--mcmc.K 32``, ``--mcmc.K=32``, ``-K 32``, ``--config F``."""
        o = cls()
        config = None
        rest = []
        i = 0
        argv = list(argv)
        while i < len(argv):
            tok = argv[i]
            if tok.startswith("--"):
                name, eq, val = tok[2:].partition("=")
            elif tok.startswith("-") and len(tok) >= 2 and not _is_number(tok):
                name, eq, val = tok[1:2], "", tok[2:]
                if val:
                    eq = "="
            else:
                rest.append(tok)
                i += 1
                continue
            if name == "config":
                config = val if eq else argv[i + 1]
                i += 1 if eq else 2
                continue
            if name in cls._SWITCHES:
                setattr(o, cls._SWITCHES[name], True)
                i += 1
                continue
            if name in cls._FLAGS:
                field, conv = cls._FLAGS[name]
                if not eq:
                    if i + 1 >= len(argv):
                        raise ValueError(f"option --{name} needs a value")
                    val = argv[i + 1]
                    i += 2
                else:
                    i += 1
                o._set(field, conv, val)
                continue
            rest.append(tok)
            i += 1
        if config is not None:
            o._read_config(config, cmdline=argv)
        o.remains = rest
        o._validate()
        return o

    def _set(self, field, conv, val):
        if field == "strategy" and val not in STRATEGIES:
            raise ValueError(f"Unknown strategy {val!r}")
        setattr(self, field, conv(val))

    def _read_config(self, path, cmdline):
        # Boost parse_config_file: "name = value" lines, [section] prefixes;
        # command-line values win (stored first, options.h:249-252)
        given = set()
        for t in cmdline:
            if t.startswith("--"):
                given.add(t[2:].partition("=")[0])
            elif t.startswith("-") and len(t) >= 2 and not _is_number(t):
                long_name = self._ALIASES.get(t[1:2])
                if long_name:
                    given.add(long_name)
        section = ""
        with open(path) as f:
            for line in f:
                line = line.split("#", 1)[0].strip()
                if not line:
                    continue
                if line.startswith("[") and line.endswith("]"):
                    section = line[1:-1].strip() + "."
                    continue
                key, _, val = line.partition("=")
                key = section + key.strip()
                val = val.strip()
                if key in given:
                    continue
                if key in self._FLAGS:
                    field, conv = self._FLAGS[key]
                    self._set(field, conv, val)
                elif key in self._SWITCHES:
                    setattr(self, self._SWITCHES[key], val.lower() in ("1", "true", "yes", "on"))

    def _validate(self):
        if self.strategy not in STRATEGIES:
            raise ValueError(f"Unknown strategy {self.strategy!r}")
        if self.precision not in ("fp64", "fp32"):
            raise ValueError("precision must be fp64 or fp32")

    def derived(self):
        """(a, alpha) after the reference's derivation (options.h:257-262)."""
        a = math.pow(self.b, -self.c) if self.a == 0.0 else self.a
        alpha = 1.0 / self.K if self.alpha == 0.0 else self.alpha
        return a, alpha

    def to_c(self):
        c = _lib.Options()
        _lib.check(_lib.lib.ammsb_options_default(ctypes.byref(c)))
        for f in ("alpha", "eta0", "eta1", "epsilon", "a", "b", "c", "K", "mini_batch_size",
                  "num_node_sample", "sampler_max_source", "max_iteration", "interval",
                  "stats_print_interval", "num_updates", "held_out_ratio", "random_seed",
                  "convergence_threshold", "device"):
            setattr(c, f, getattr(self, f))
        c.strategy = STRATEGIES[self.strategy]
        c.precision = _lib.FP64 if self.precision == "fp64" else _lib.FP32
        c.verbose = 1 if self.verbose else 0
        return c


def _is_number(tok):
    try:
        float(tok)
        return True
    except ValueError:
        return False
