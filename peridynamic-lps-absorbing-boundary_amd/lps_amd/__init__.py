"""MI355X-native LPS peridynamics time loop with least-squares absorbing boundaries."""
