%--- phisical properties
  1
  1   FLUID  0.01   1.00  0.01  1.00 
%--- boundary conditions
   4 
   1    WALL      0.0   0.0   0.0   0.0
   2    WALLFLUX  0.0   0.0   0.0   1.00
   3    WALLFLUX  0.0   0.0   0.0   0.0
   4    WALLFLUX  0.0   0.0   0.0   0.0
%---- initial conditions
   1
   1   0.0  0.0  0.0  0.0
