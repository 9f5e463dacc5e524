"""Minimal configuration object with the reading interface of the reference's
``Inifile`` (``pyfr/inifile.py:31-161``): ``get``, ``getint``, ``getfloat``,
``getbool``, ``hasopt``, ``set``, ``items``, ``items_as`` and ``tostr``."""
import configparser
import io
import re


class Inifile:
    def __init__(self, inistr=None):
        self._cp = configparser.ConfigParser(inline_comment_prefixes=[';'])
        self._cp.optionxform = str
        if inistr:
            self._cp.read_string(inistr)

    @classmethod
    def from_dict(cls, d):
        cfg = cls()
        for sect, opts in d.items():
            for k, v in opts.items():
                cfg.set(sect, k, v)
        return cfg

    def set(self, sect, opt, value):
        if not self._cp.has_section(sect):
            self._cp.add_section(sect)
        self._cp.set(sect, opt, str(value))

    def hasopt(self, sect, opt):
        return self._cp.has_option(sect, opt)

    def tostr(self):
        """The configuration as ini text (what solution files record)."""
        buf = io.StringIO()
        self._cp.write(buf)
        return buf.getvalue()

    _NONE = object()

    def get(self, sect, opt, default=_NONE):
        if self._cp.has_option(sect, opt):
            return self._cp.get(sect, opt)
        if default is self._NONE:
            raise KeyError(f'[{sect}] {opt} missing')
        return default if default is None else str(default)

    def getint(self, sect, opt, default=_NONE):
        v = self.get(sect, opt, default)
        return None if v is None else int(v)

    def getfloat(self, sect, opt, default=_NONE):
        v = self.get(sect, opt, default)
        return None if v is None else float(v)

    def getbool(self, sect, opt, default=_NONE):
        v = self.get(sect, opt, default)
        return v if isinstance(v, bool) else str(v).lower() in \
            ('1', 'true', 'yes', 'on')

    def getexpr(self, sect, opt, default=_NONE):
        """Expression with integer literals promoted to floats
        (``pyfr/inifile.py:14-25``, ``:87-88``)."""
        expr = self.get(sect, opt, default)
        if not re.match(r'[A-Za-z0-9 \t\n\r.,+\-*/%()<>=&|!?:_]+$', expr):
            raise ValueError(f'Invalid characters in expression {expr!r}')
        def fix(m):
            tok = m.group(0)
            return tok + '.0' if tok.isdigit() else tok
        return re.sub(r'(?<![\w.])\d+\.?\d*(?:[eE][+-]?\d+)?', fix, expr)

    def items(self, sect):
        if not self._cp.has_section(sect):
            return {}
        return dict(self._cp.items(sect))

    def items_as(self, sect, type_):
        out = {}
        for k, v in self.items(sect).items():
            try:
                out[k] = type_(v)
            except ValueError:
                pass
        return out
