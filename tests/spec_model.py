"""A second, independent statement of the step in plain Python (small grids).

Written from the model description, not from the oracle's code: agents are
objects with (x, y) coordinates, cells are dictionaries of agent lists, and
the deterministic choices are restated here:
  * RNG: Philox4x32-10 blocks (the primitive is taken from the oracle, whose
    Philox is pinned by published known answers and by rocRAND), key =
    (seed low, seed high), counter = (agent id or cell, step, stream, block),
    streams 0 init cell, 1 init agent, 2 move, 3 shuffle; bounded draw
    floor(word * n / 2^32);
  * agents of a cell listed by id, Fisher-Yates from the last position with
    draw m = k-1-i taken from word m%4 of block m//4;
  * next state: cells ascending, survivors in acting order, child right after
    its parent.
Model rules (NetLogo pphpc.nlogo:63-159, Java Cell/Sheep/Wolf/AbstractAgent):
move with energy -1 and death below 1, grass countdown -1, then per cell in
shuffled order: sheep eat live grass (countdown := R, +gain), wolves eat the
first uneaten sheep of the list (+gain), reproduction splits energy
(child floor(e/2)) when e > threshold and the draw is below the probability.
Energy limit: an agent whose energy after its action exceeds 2^24 - 1 makes
the step fail with -9 (the GPU's packed energy field is 24 bits wide).
"""
from oracle.pp_oracle import philox

SHEEP, WOLF = 0, 1
ENERGY_MAX = (1 << 24) - 1
ENERGY_OVERFLOW = -9


def draw(key, a, step, stream, block):
    return philox((a & 0xFFFFFFFF, step, stream, block), key)


def bounded(word, n):
    return (word * n) >> 32


class Agent:
    __slots__ = ("kind", "x", "y", "energy", "dice", "eaten", "child")

    def __init__(self, kind, x, y, energy):
        self.kind, self.x, self.y, self.energy = kind, x, y, energy
        self.dice = None
        self.eaten = False
        self.child = None


class SpecModel:
    def __init__(self, p, seed):
        self.p = p
        self.key = (seed & 0xFFFFFFFF, seed >> 32)
        self.t = 0
        gx, gy = p["grid_x"], p["grid_y"]
        self.countdown = {}
        for y in range(gy):
            for x in range(gx):
                c = y * gx + x
                w = draw(self.key, c, 0, 0, 0)
                alive = bounded(w[0], 2) == 1
                self.countdown[(x, y)] = 0 if alive else \
                    1 + bounded(w[1], p["grass_restart"])
        self.agents = []
        n0 = p["init_sheep"] + p["init_wolves"]
        for i in range(n0):
            kind = SHEEP if i < p["init_sheep"] else WOLF
            gain = p["sheep_gain_from_food"] if kind == SHEEP else \
                p["wolves_gain_from_food"]
            w = draw(self.key, i, 0, 1, 0)
            self.agents.append(Agent(kind, bounded(w[0], gx),
                                     bounded(w[1], gy),
                                     1 + bounded(w[2], 2 * gain)))

    def step(self):
        p, gx, gy = self.p, self.p["grid_x"], self.p["grid_y"]
        moves = {0: (0, 0), 1: (1, 0), 2: (-1, 0), 3: (0, 1), 4: (0, -1)}
        cells = {}
        for i, a in enumerate(self.agents):
            w = draw(self.key, i, self.t, 2, 0)
            dx, dy = moves[bounded(w[0], 5)]
            a.dice = bounded(w[1], 100)
            a.x, a.y = (a.x + dx) % gx, (a.y + dy) % gy
            a.energy -= 1
            a.eaten, a.child = False, None
            if a.energy >= 1:
                cells.setdefault((a.x, a.y), []).append(a)  # id order
        for xy in self.countdown:
            if self.countdown[xy] > 0:
                self.countdown[xy] -= 1
        new_agents = []
        over = False
        for (x, y) in sorted(cells, key=lambda xy: xy[1] * gx + xy[0]):
            lst = cells[(x, y)]
            c = y * gx + x
            k = len(lst)
            words = []
            for b in range((k + 2) // 4 + 1):
                words.extend(draw(self.key, c, self.t, 3, b))
            for m, i in enumerate(range(k - 1, 0, -1)):
                j = bounded(words[m], i + 1)
                lst[i], lst[j] = lst[j], lst[i]
            for a in lst:
                if a.eaten:
                    continue
                if a.kind == SHEEP:
                    if self.countdown[(x, y)] == 0:
                        self.countdown[(x, y)] = p["grass_restart"]
                        a.energy += p["sheep_gain_from_food"]
                    thr, prob = (p["sheep_reproduce_threshold"],
                                 p["sheep_reproduce_prob"])
                else:
                    prey = next((b for b in lst
                                 if b.kind == SHEEP and not b.eaten), None)
                    if prey is not None:
                        prey.eaten = True
                        a.energy += p["wolves_gain_from_food"]
                    thr, prob = (p["wolves_reproduce_threshold"],
                                 p["wolves_reproduce_prob"])
                if a.energy > thr and a.dice < prob:
                    child = Agent(a.kind, x, y, a.energy // 2)
                    a.energy -= child.energy
                    a.child = child
                if a.energy > ENERGY_MAX:
                    over = True
            for a in lst:
                if not a.eaten:
                    new_agents.append(a)
                if a.child is not None:
                    new_agents.append(a.child)
        if over:
            return ENERGY_OVERFLOW
        self.agents = new_agents
        self.t += 1
        return 0

    def state(self):
        gx = self.p["grid_x"]
        cell = [a.y * gx + a.x for a in self.agents]
        energy = [a.energy for a in self.agents]
        kind = [a.kind for a in self.agents]
        grass = [self.countdown[(c % gx, c // gx)]
                 for c in range(gx * self.p["grid_y"])]
        return cell, energy, kind, grass

    def stats(self):
        s = [a for a in self.agents if a.kind == SHEEP]
        w = [a for a in self.agents if a.kind == WOLF]
        g = list(self.countdown.values())
        return (len(s), len(w), sum(1 for v in g if v == 0),
                sum(a.energy for a in s), sum(a.energy for a in w), sum(g), 0)
