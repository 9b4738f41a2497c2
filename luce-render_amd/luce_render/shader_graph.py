"""Material node graphs compiled to programs for the kernels
(shader_graph.lucb; the interpreter is csrc/render_program.h).

A material's graph is text, one statement a line (or separated by ';'):

    node <id> <kind> [<setting>]      the setting is the rest of the line
    input <id> <input> <x> [<y> <z>]  a constant for an unlinked input
    link <id> <input> <source id> <output>

The `openpbr_surface` node is the terminal: its linked inputs override the
material's parameters. A program is a flat list of u32 words: an opcode, then
its operands and output stack slots. An operand is a constant's f32 bits or a
stack slot tagged in a NaN (0x7FC10000 | slot); colors and vectors take three
words. A bump compiles its height's nodes three times: at the point and a
footprint away along each of two tangents (uv and position ops carry the
tangent and the footprint's operand).
"""
import numpy as np

from ._lib import INVALID, LuceError

PROGRAM_STACK = 64
SLOT_TAG = 0x7FC10000

OP_END, OP_UV, OP_POSITION, OP_VALUE, OP_COLOR, OP_IMAGE, OP_CHECKER, OP_NOISE, OP_MATH, OP_MIX = range(10)
OP_MAP_RANGE, OP_CLAMP, OP_SEPARATE, OP_COMBINE, OP_RAMP, OP_NORMAL_MAP, OP_OUTPUT, OP_BUMP = range(10, 18)
OP_FRESNEL, OP_LAYER_WEIGHT = 18, 19

MATH_OPERATIONS = ["add", "subtract", "multiply", "divide", "power", "minimum", "maximum", "absolute", "sine",
                   "cosine", "floor", "fraction"]


def _one(name, value):
    return (name, 1, (value, 0.0, 0.0))


def _three(name, x, y, z):
    return (name, 3, (x, y, z))


# Each kind's opcode, inputs and outputs (socket name, width, default), as the
# reference's kind_of lists them.
KINDS = {
    "uv": (OP_UV, [], [_three("uv", 0, 0, 0)]),
    "position": (OP_POSITION, [], [_three("position", 0, 0, 0)]),
    "value": (OP_VALUE, [_one("value", 0.0)], [_one("value", 0.0)]),
    "color": (OP_COLOR, [_three("color", 0.8, 0.8, 0.8)], [_three("color", 0, 0, 0)]),
    "image": (OP_IMAGE, [_three("vector", 0, 0, 0)], [_three("color", 0, 0, 0), _one("alpha", 0.0)]),
    "checker": (OP_CHECKER, [_three("vector", 0, 0, 0), _three("color1", 0.8, 0.8, 0.8), _three("color2", 0.2, 0.2, 0.2),
                             _one("scale", 5.0)], [_three("color", 0, 0, 0), _one("fac", 0.0)]),
    "noise": (OP_NOISE, [_three("vector", 0, 0, 0), _one("scale", 5.0), _one("detail", 2.0), _one("roughness", 0.5)],
              [_one("fac", 0.0), _three("color", 0, 0, 0)]),
    "math": (OP_MATH, [_one("a", 0.5), _one("b", 0.5)], [_one("value", 0.0)]),
    "mix": (OP_MIX, [_one("fac", 0.5), _three("a", 0.5, 0.5, 0.5), _three("b", 0.5, 0.5, 0.5)],
            [_three("color", 0, 0, 0)]),
    "map_range": (OP_MAP_RANGE, [_one("value", 0.5), _one("from_min", 0.0), _one("from_max", 1.0),
                                 _one("to_min", 0.0), _one("to_max", 1.0)], [_one("value", 0.0)]),
    "clamp": (OP_CLAMP, [_one("value", 0.5), _one("min", 0.0), _one("max", 1.0)], [_one("value", 0.0)]),
    "separate": (OP_SEPARATE, [_three("color", 0.8, 0.8, 0.8)], [_one("r", 0.0), _one("g", 0.0), _one("b", 0.0)]),
    "combine": (OP_COMBINE, [_one("r", 0.0), _one("g", 0.0), _one("b", 0.0)], [_three("color", 0, 0, 0)]),
    "ramp": (OP_RAMP, [_one("fac", 0.5), _three("color1", 0, 0, 0), _three("color2", 1, 1, 1), _one("position1", 0.0),
                       _one("position2", 1.0)], [_three("color", 0, 0, 0)]),
    "normal_map": (OP_NORMAL_MAP, [_three("color", 0.5, 0.5, 1.0), _one("strength", 1.0)],
                   [_three("normal", 0, 0, 0)]),
    # an unlinked normal is zero, which the kernels read as the shading normal
    "bump": (OP_BUMP, [_one("strength", 1.0), _one("distance", 1.0), _one("filter_width", 0.1), _one("height", 1.0),
                       _three("normal", 0, 0, 0)], [_three("normal", 0, 0, 0)]),
    "fresnel": (OP_FRESNEL, [_one("ior", 1.5), _three("normal", 0, 0, 0)], [_one("fac", 0.0)]),
    "layer_weight": (OP_LAYER_WEIGHT, [_one("blend", 0.5), _three("normal", 0, 0, 0)],
                     [_one("fresnel", 0.0), _one("facing", 0.0)]),
}

# The OpenPBR parameters a graph may drive: (register, width, mask bit).
SURFACE_REGISTERS = {
    "base_color": (0, 3, 0), "base_weight": (3, 1, 1), "base_metalness": (4, 1, 2),
    "base_diffuse_roughness": (5, 1, 3), "specular_weight": (6, 1, 4), "specular_roughness": (7, 1, 5),
    "coat_weight": (8, 1, 6), "coat_roughness": (9, 1, 7), "transmission_weight": (10, 1, 8),
    "geometry_normal": (11, 3, 9), "specular_color": (14, 3, 10), "transmission_color": (17, 3, 11),
    "coat_color": (20, 3, 12),
}


def constant_bits(value):
    """A constant operand: its f32 bits (non-finite values become 0)."""
    with np.errstate(over="ignore", invalid="ignore"):
        v = np.float32(value)
    if not np.isfinite(v) or v > 3.0e38 or v < -3.0e38:
        v = np.float32(0.0)
    return int(np.array(v, np.float32).view(np.uint32))


def _number(text):
    if not text:
        return 0.0
    try:
        return float(text)
    except ValueError:
        raise LuceError(INVALID, f"a node graph's number {text!r} does not parse")


class _Node:
    def __init__(self, ident, kind, setting):
        self.id, self.kind, self.setting = ident, kind, setting
        self.emitted = False
        self.slots = []


class _Compiler:
    def __init__(self, resolve_image):
        self.nodes, self.inputs, self.words = [], [], []
        self.stack = 0
        self.resolve_image = resolve_image
        self.uv_slot = self.position_slot = -1
        self.axis = 0
        self.pixels = 0

    def find(self, ident):
        for at, n in enumerate(self.nodes):
            if n.id == ident:
                return at
        return -1

    def input_of(self, node, name):
        for at, e in enumerate(self.inputs):
            if e["node"] == node and e["name"] == name:
                return at
        return -1

    def allocate(self, width):
        slot = self.stack
        if self.stack + width > PROGRAM_STACK:
            raise LuceError(INVALID, "a material's node graph needs more than 64 values at once")
        self.stack += width
        return slot

    def implicit(self, position):
        held = self.position_slot if position else self.uv_slot
        if held >= 0:
            return held
        slot = self.allocate(3)
        self.words += [OP_POSITION if position else OP_UV, self.axis, self.pixels, slot, slot + 1, slot + 2]
        if position:
            self.position_slot = slot
        else:
            self.uv_slot = slot
        return slot

    def operands(self, node, kind_name, socket, out):
        name, width, default = socket
        at = self.input_of(node, name)
        if at >= 0 and self.inputs[at]["source"] >= 0:
            e = self.inputs[at]
            slot = self.output_slot(e["source"], e["output"], width)
            out += [SLOT_TAG | (slot + k) for k in range(width)]
            return
        if at < 0 and name == "vector":
            slot = self.implicit(kind_name == "noise")
            out += [SLOT_TAG | (slot + k) for k in range(width)]
            return
        values = self.inputs[at]["values"] if at >= 0 else default
        out += [constant_bits(values[k]) for k in range(width)]

    def output_slot(self, source, name, width):
        self.emit(source, 0)
        _, _, outputs = KINDS[self.nodes[source].kind]
        for at, (oname, owidth, _) in enumerate(outputs):
            if oname != name:
                continue
            slot = self.nodes[source].slots[at]
            if owidth == width:
                return slot
            # widths differ: a value spread to a color, or a color's first
            spread = self.allocate(width)
            if width == 3:
                self.words += [OP_COLOR] + [SLOT_TAG | slot] * 3 + [spread, spread + 1, spread + 2]
            else:
                self.words += [OP_VALUE, SLOT_TAG | slot, spread]
            return spread
        raise LuceError(INVALID, "a node graph links an output its node does not have")

    def offset_slot(self, source, name, axis, pixels):
        saved = [(n.emitted, n.slots) for n in self.nodes]
        for n in self.nodes:
            n.emitted, n.slots = False, []
        uv, position = self.uv_slot, self.position_slot
        self.uv_slot = self.position_slot = -1
        self.axis, self.pixels = axis, pixels
        try:
            return self.output_slot(source, name, 1)
        finally:
            for n, (emitted, slots) in zip(self.nodes, saved):
                n.emitted, n.slots = emitted, slots
            self.uv_slot, self.position_slot = uv, position
            self.axis = self.pixels = 0

    def heights(self, node, socket, out):
        if self.axis != 0:
            raise LuceError(INVALID, "a Bump's height cannot come from another Bump")
        self.operands(node, "bump", socket, out)
        at = self.input_of(node, socket[0])
        center = out[-1]
        if at < 0 or self.inputs[at]["source"] < 0:
            out += [center, center]
            return
        e = self.inputs[at]
        pixels = out[3]  # the filter width's operand: after the opcode, strength and distance
        for axis in (1, 2):
            out.append(SLOT_TAG | self.offset_slot(e["source"], e["output"], axis, pixels))

    def emit(self, index, depth):
        if depth > 256:
            raise LuceError(INVALID, "a material's node graph loops")
        node = self.nodes[index]
        if node.emitted:
            return
        if node.kind not in KINDS:
            raise LuceError(INVALID, "a material's node graph has a node of an unknown kind")
        opcode, inputs, outputs = KINDS[node.kind]
        for e in self.inputs:
            if e["node"] == index and e["source"] >= 0:
                self.emit(e["source"], depth + 1)
        head = [opcode]
        if opcode == OP_IMAGE:
            setting, color = node.setting, True
            if len(setting) > 5 and setting.endswith(" data"):
                setting, color = setting[:-5], False
            head += [int(self.resolve_image(setting, color)), 1 if color else 0]
        if opcode == OP_MATH:
            head.append(MATH_OPERATIONS.index(node.setting) if node.setting in MATH_OPERATIONS else 0)
        if opcode in (OP_UV, OP_POSITION):
            head += [self.axis, self.pixels]
        for socket in inputs:
            if opcode == OP_BUMP and socket[0] == "height":
                self.heights(index, socket, head)
            else:
                self.operands(index, node.kind, socket, head)
        self.words += head
        for _, width, _ in outputs:
            slot = self.allocate(width)
            node.slots.append(slot)
            self.words += [slot + k for k in range(width)]
        node.emitted = True


def _tokens(line):
    return line.split()


def compile_graph(text, resolve_image=None):
    """A graph's program (a list of u32 words ending in OP_END), or None when
    the graph drives no parameter. `resolve_image(file, color)` gives an image
    node's texture slot."""
    def no_images(name, color):
        raise LuceError(INVALID, f"a node graph reads image {name!r} but the material has no such image")

    c = _Compiler(resolve_image or no_images)
    lines = [ln for part in text.split("\n") for ln in part.split(";")]
    for ln in lines:  # nodes first, so links may name nodes defined later
        t = _tokens(ln)
        if t and t[0] == "node":
            setting = ln.split(None, 3)[3] if len(t) > 3 else ""
            c.nodes.append(_Node(t[1] if len(t) > 1 else "", t[2] if len(t) > 2 else "", setting.strip()))
    for ln in lines:
        t = _tokens(ln)
        if not t or t[0] not in ("input", "link"):
            continue
        node = c.find(t[1] if len(t) > 1 else "")
        if node < 0:
            raise LuceError(INVALID, "a node graph sets an input of a node it lacks")
        entry = {"node": node, "name": t[2] if len(t) > 2 else "", "values": [0.0, 0.0, 0.0], "source": -1,
                 "output": ""}
        if t[0] == "input":
            entry["values"] = [_number(t[3 + k] if len(t) > 3 + k else "") for k in range(3)]
            if len(t) <= 4:
                entry["values"][1] = entry["values"][2] = entry["values"][0]
        else:
            entry["source"] = c.find(t[3] if len(t) > 3 else "")
            if entry["source"] < 0:
                raise LuceError(INVALID, "a node graph links from a node it lacks")
            entry["output"] = t[4] if len(t) > 4 else ""
        c.inputs.append(entry)
    surface = -1
    for at, n in enumerate(c.nodes):
        if n.kind == "openpbr_surface":
            surface = at
    if surface < 0:
        return None
    driven = False
    for e in list(c.inputs):
        if e["node"] != surface or e["source"] < 0 or e["name"] not in SURFACE_REGISTERS:
            continue
        register, width, bit = SURFACE_REGISTERS[e["name"]]
        slot = c.output_slot(e["source"], e["output"], width)
        c.words += [OP_OUTPUT, register, width, bit] + [SLOT_TAG | (slot + k) for k in range(width)]
        driven = True
    if not driven:
        return None
    c.words.append(OP_END)
    return c.words
