"""mini::temporal surface: the ODE System contract and the explicit solvers.

Mirrors include/mini/temporal/ode.hpp:17-134 and temporal/rk.hpp:10-57.  The
operation order of every update is the reference's, so a column advanced by
``RungeKutta(3).update`` through any System equals, bit for bit, what the
reference computes through the same System.
"""
from __future__ import annotations

import abc

import numpy as np


class System(abc.ABC):
    """temporal/ode.hpp:17-48: four pure virtuals."""

    @abc.abstractmethod
    def set_time(self, t_curr: float) -> None: ...

    @abc.abstractmethod
    def set_solution_column(self, column: np.ndarray) -> None: ...

    @abc.abstractmethod
    def get_solution_column(self) -> np.ndarray: ...

    @abc.abstractmethod
    def get_residual_column(self) -> np.ndarray: ...


class Constant(System):
    """temporal/ode.hpp:55-85: dU/dt = A U."""

    def __init__(self, a: np.ndarray):
        assert a.shape[0] == a.shape[1]
        self._a = a
        self._u = None

    def set_time(self, t_curr):
        pass

    def set_solution_column(self, column):
        assert column.shape[0] == self._a.shape[1]
        self._u = column.copy()

    def get_solution_column(self):
        return self._u.copy()

    def get_residual_column(self):
        return self._a @ self._u


class Euler:
    """temporal/ode.hpp:112-134."""

    @staticmethod
    def next_solution(system: System, t_curr: float, dt: float) -> np.ndarray:
        system.set_time(t_curr)
        u_next = system.get_solution_column()
        residual = system.get_residual_column()
        residual *= dt
        u_next += residual
        return u_next

    def update(self, system: System, t_curr: float, dt: float) -> None:
        system.set_solution_column(self.next_solution(system, t_curr, dt))


class RungeKutta:
    """temporal/rk.hpp:10-57 (Shu-Osher SSP-RK of order 1, 2 or 3)."""

    def __init__(self, orders: int):
        assert orders in (1, 2, 3)
        self.orders = orders
        self._euler = Euler()

    def update(self, system: System, t_curr: float, dt: float) -> None:
        if self.orders == 1:
            self._euler.update(system, t_curr, dt)
        elif self.orders == 2:
            u_curr = system.get_solution_column()
            self._euler.update(system, t_curr, dt)
            u_next = self._euler.next_solution(system, t_curr + dt, dt)
            u_next += u_curr
            u_next *= 0.5
            system.set_solution_column(u_next)
        else:
            u_curr = system.get_solution_column()
            self._euler.update(system, t_curr, dt)
            u_next = u_curr.copy()
            u_next *= 3
            u_next += self._euler.next_solution(system, t_curr + dt, dt)
            u_next /= 4
            system.set_solution_column(u_next)
            u_next = self._euler.next_solution(system, t_curr + dt / 2, dt)
            u_next *= 2
            u_next += u_curr
            u_next /= 3
            system.set_solution_column(u_next)
