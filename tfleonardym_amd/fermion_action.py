"""Rational pseudofermion action, drop-in for the reference's
fermion_action.n_flavor (fermion_action.py:7-93).

Same class, constructor, attributes and methods.  The pseudofermions, the
energy, the rational checks and the force run on the device: the multishift
solves through the given solver (liblym CG-M / MR), Q x through the Wilson
kernel, and the weighted sum of Lie derivatives over every pole of every
force approximation through the force kernel, one launch for all of them
when the operator is a DiracWilsonOperator.
"""
import numpy as np
import torch

from . import _lib
from . import algebra_utils
from . import dirac
from .tensors import to_device


class n_flavor:
    def __init__(self, force_operators, metropolis_operators, heatbath_operators, solver, dirac_operator):
        self.force_operators = force_operators
        self.metropolis_operators = metropolis_operators
        self.heatbath_operators = heatbath_operators
        self.solver = solver
        self.dirac_operator = dirac_operator
        self.square_dirac_operator = dirac.SquareOperator(dirac_operator)
        self.pseudofermions = None

    def initialize_pseudofermions(self, random_vectors):
        """phi_k = R_hb,k(A) eta_k (fermion_action.py:22-25)."""
        self.pseudofermions = [operator.apply(vector)
                               for vector, operator in zip(random_vectors, self.heatbath_operators)]

    def set_gauge_configuration(self, cfg):
        """fermion_action.py:27-37."""
        self.dirac_operator.set_gauge_configuration(cfg)
        self.square_dirac_operator = dirac.SquareOperator(self.dirac_operator)
        for approximation in self.force_operators:
            approximation.set_gauge_configuration(cfg)
        for approximation in self.metropolis_operators:
            approximation.set_gauge_configuration(cfg)
        for approximation in self.heatbath_operators:
            approximation.set_gauge_configuration(cfg)

    def energy(self):
        """Re sum_k <phi_k, R_met,k(A) phi_k> (fermion_action.py:39-44)."""
        energy = 0
        for vector, operator in zip(self.pseudofermions, self.metropolis_operators):
            energy += algebra_utils.dot(vector, operator.apply(vector))
        return float(np.real(energy))

    def number_of_flavors(self):
        """2 log prod_k r_met,k(0.5) / log 2 (fermion_action.py:46-52)."""
        result = 1
        for approximation in self.metropolis_operators:
            result *= approximation.evaluate(0.5)
        return 2 * np.log(result) / np.log(2.)

    def check_rational_approximations(self, vector):
        """sum_k |R_hb R_met R_hb v - v|^2 (fermion_action.py:54-62)."""
        v = to_device(vector)
        result = 0.
        for metropolis, heatbath in zip(self.metropolis_operators, self.heatbath_operators):
            third = heatbath.apply(metropolis.apply(heatbath.apply(v)))
            algebra_utils.axpby(-1.0, v, 1.0, third)
            result += algebra_utils.dot(third, third)
        return result

    def force(self):
        """sum_k sum_i alpha_ki [L(x_ki, Q x_ki) + L(Q x_ki, x_ki)] with
        (A + beta_ki) x_ki = phi_k (fermion_action.py:64-93)."""
        Xs = [self.solver.solve(force_operator.operator, pseudofermion, force_operator.betas)
              for force_operator, pseudofermion in zip(self.force_operators, self.pseudofermions)]
        Ys = [[self.dirac_operator.apply(solution) for solution in x] for x in Xs]
        pairs, weights = [], []
        for X, Y, approximation in zip(Xs, Ys, self.force_operators):
            for x, y, weight in zip(X, Y, approximation.alphas):
                pairs.append((x, y))
                weights.append(weight)
        op = self.dirac_operator
        if isinstance(op, dirac.DiracWilsonOperator):
            return op.force_sum(pairs, weights, symmetric=True)
        # any other operator with a lie_derivative: the reference's loop
        force = torch.zeros((op.number_of_dimensions, op.number_of_colors, op.number_of_colors, op.volume),
                            dtype=torch.complex128, device=_lib.require_device())
        for (x, y), weight in zip(pairs, weights):
            algebra_utils.axpby(weight, op.lie_derivative(x, y), 1.0, force)
            algebra_utils.axpby(weight, op.lie_derivative(y, x), 1.0, force)
        return force
